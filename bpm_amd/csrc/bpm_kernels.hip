// Split-step beam propagation on gfx950: fused row/column FFT passes.
//
// One propagation step z-1 -> z of the reference loop (bpm/bpm/bpm_3d.py:283-382;
// bpm/bpmclass/_bpm3d_ocl.py:100-126) is
//     u_z = screen_z( IFFT2( H . FFT2(u_{z-1}) ) )
// and is executed as two passes over the (Ny,Nx) complex64 plane that stays
// in HBM between them, holding F_x u (the plane transformed along x only):
//   column pass (one work-group per tile of COLS columns):
//       FFT_y -> multiply by H (table or on-the-fly formula) -> IFFT_y
//   row pass (one work-group per LINES rows):
//       IFFT_x -> 1/(NxNy) -> phase screen dn[z] and taper -> outputs
//       (field/intensity slice, running sum of |u|^2) -> FFT_x
// so each step reads and writes the plane twice and reads dn[z] once.  The
// default mean_mode "weighted" of bpm_3d needs the k-space reduction of
// bpm_3d.py:302-303 before H, so its column pass is split in two (A: FFT_y
// and per-column partial sums; B: H(n0+dn0) and IFFT_y).
#include "common.hpp"
#include "fft.hpp"
#include "bpm_hip.h"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

using namespace bpmfft;

// ----------------------------------------------------------------------------
// error reporting
// ----------------------------------------------------------------------------
static thread_local std::string g_err;

int bpmh_set_error(int code, const char* msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                        \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess) {                                                  \
      char b_[256];                                                          \
      snprintf(b_, sizeof b_, "%s: %s", #expr, hipGetErrorString(e_));       \
      return bpmh_set_error(BPMH_ERR_HIP, b_);                               \
    }                                                                        \
  } while (0)

extern "C" const char* bpmh_last_error(void) { return g_err.c_str(); }
extern "C" int bpmh_version(void) { return BPMH_VERSION; }

// ----------------------------------------------------------------------------
// per-size launch configurations: variant V in {0,1,2} scales the default
// number of lines (rows) or columns per work-group by 1/2, 1, 2, clamped to
// 1..16 lines, 1024 threads and 128 KiB of LDS (gfx950 has 160 KiB per CU);
// the autotuner picks V per size.
// ----------------------------------------------------------------------------
constexpr int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
constexpr int pow2floor(int v) { return v < 2 ? 1 : 2 * pow2floor(v / 2); }
constexpr int scale_v(int base, int V) { return V == 0 ? base / 2 : (V == 1 ? base : base * 2); }

// Elements per thread of a line.  Rows: 4 up to N = 1024 (twice the
// wavefronts of E = 8 for the same data, so loads, transforms and stores of
// different wavefronts overlap better; measured 5-13 % faster per z-step at
// 64^2..1024^2, batch 1), 8 above.  Columns: 8 (4-column tiles of 8-element
// threads; E = 4 measured slower at 1024^2, also with 4-column tiles of
// 1024 threads, and at 256^2..512^2 in batches).
constexpr int row_e(int n) { return n <= 1024 ? 4 : 8; }
constexpr int col_e(int n) { return n <= 32 ? 4 : 8; }
// The other radix schedule of a size, which the autotuner times against the
// default: 8 elements per thread instead of 4 (when the line still holds two
// radix-8 stages, N >= 64), or 4 instead of 8.  ealt selects it.
constexpr int alt_e(int e, int n) { return e == 4 ? (n >= 64 ? 8 : 4) : 4; }
constexpr int row_e_of(int n, int ealt) { return ealt ? alt_e(row_e(n), n) : row_e(n); }
constexpr int col_e_of(int n, int ealt) { return ealt ? alt_e(col_e(n), n) : col_e(n); }
// Variant numbers: row variant = V + 3*ealt (V = 0, 1, 2), column variant =
// V + 5*ealt (V = 0..4).
constexpr int kRowShapes = 3, kColShapes = 5;
#define BPMH_MAX_GROUPS 16


template <int N, int E_>
struct Base {
  static constexpr int E = E_;
  static constexpr int T = N / E;
  // LDS (layouts of fft.hpp): a row line is N float2 in the swizzled
  // exchange layout (8 more when lines of fewer than 16 threads share a
  // 16-lane group, so those lines start in different banks).  Column tiles
  // of 4 or more columns use the swizzled layout for C = COLS lines per
  // 16-lane group, lines N + 16/COLS apart; narrower tiles keep lines padded
  // one per 32 and 9 float2 longer (the swizzle measured slower there).
  // COL_LINE, the padded length, bounds both.
  static constexpr int ROW_LINE = N + (T < 16 ? 8 : 0);
  static constexpr int COL_SH = 5;
  static constexpr int COL_LINE = N + (N >> COL_SH) + 9;  // float2 per column line
  static constexpr int LDS_LINE = COL_LINE * 8;        // bytes per column line
};

template <int N, int V, int E_ = row_e(N)>
struct RowCfg : Base<N, E_> {
  using Base<N, E_>::T;
  static constexpr int L0 = clampi(256 / T, 1, 16);
  // at least one full wavefront per work-group (the cross-lane reductions
  // use all 64 lanes), at most 16 lines and 1024 threads
  static constexpr int LMIN = 64 / T < 1 ? 1 : 64 / T;
  static constexpr int LINES =
      clampi(clampi(scale_v(L0, V), LMIN, 16), 1, (1024 / T) < 1 ? 1 : 1024 / T);
  static constexpr int THREADS = LINES * T;
};

template <int N, int V, int E_ = col_e(N)>
struct ColCfg : Base<N, E_> {
  using Base<N, E_>::T;
  using Base<N, E_>::LDS_LINE;
  static constexpr int C0 = N >= 4096 ? 1 : clampi(512 / T, 2, 16);
  static constexpr int CMAX0 = 131072 / LDS_LINE < 1 ? 1 : 131072 / LDS_LINE;
  // <= 512 threads with 8 elements per thread (1024 spilled), <= 1024 with 4
  static constexpr int TMAX = (E_ <= 4 ? 1024 : 512) / T < 1 ? 1 : (E_ <= 4 ? 1024 : 512) / T;
  static constexpr int CMAX = pow2floor(clampi(CMAX0 < TMAX ? CMAX0 : TMAX, 1, 16));
  static constexpr int CMIN = 64 / T < 1 ? 1 : 64 / T;  // >= one full wavefront
  // V = 3: the work-group of V = 0 with P = 2 adjacent columns per thread
  // (tiles twice as wide), so a row of a column pair is one 16-byte load and
  // store.  V = 4: the same with each column's loads and stores issued
  // separately (8-byte accesses), the first column's loads first, so its
  // transform can start while the second column's loads are in flight.
  // Used for columns of 256-1024 points (where they were measured);
  // elsewhere they are V = 0.
  static constexpr bool PAIRED = (V == 3 || V == 4) && N >= 256 && N <= 1024;
  static constexpr int P = PAIRED ? 2 : 1;
  static constexpr bool SPLIT = PAIRED && V == 4;
  static constexpr int LC = clampi(scale_v(C0, V >= 3 ? 0 : V), CMIN, CMAX);  // lane columns
  static constexpr int COLS = LC * P;
  static constexpr int THREADS = LC * T;
};

// ----------------------------------------------------------------------------
// device helpers
// ----------------------------------------------------------------------------
BPMH_DEV float2 load_dn(const void* dn, int kind, size_t idx) {
  if (kind == BPMH_DN_F32) return make_float2(((const float*)dn)[idx], 0.f);
  if (kind == BPMH_DN_F16) return make_float2(__half2float(((const __half*)dn)[idx]), 0.f);
  return ((const float2*)dn)[idx];
}

// FFTFREQ(i,N,d) of bpm_3d_kernels.cl:6 in float32
BPMH_DEV float fftfreq_f32(int i, int N, float d) {
  const int half = (N - 1) / 2 + 1;
  const float num = (i < half) ? (float)i : (float)((i - half) - N / 2);
  return f_div(f_div(num, (float)N), d);
}

BPMH_DEV float kcomp2_f32(int i, int N, float d) {
  const float k = f_mul(2.f * BPMH_PI_F, fftfreq_f32(i, N, d));
  return f_mul(k, k);
}

// H for one (kx^2, ky^2) with a = ((n*n)*k0)*k0, as compute_propagator
// (bpm_3d_kernels.cl:419-430): (tmp>=0) * exp(-i dz sqrt(tmp)).
BPMH_DEV float2 prop_f32(float kx2, float ky2, float a, float dz) {
  const float tmp = f_sub(f_sub(a, kx2), ky2);
  const float h0 = tmp < 0.f ? 0.f : __fsqrt_rn(tmp);
  const float2 cs = sincos_f32_via_f64(f_mul(-dz, h0));
  return tmp >= 0.f ? cs : make_float2(0.f, 0.f);
}

BPMH_DEV float n_term_f32(float n, float k0) { return f_mul(f_mul(f_mul(n, n), k0), k0); }

// prop_f32 (the same operations, so the same H) that also returns
// beta = sqrt(a - kx^2 - ky^2), 0 where the component does not propagate
// (H = 0 there, and its derivative with respect to dn0 is taken as 0).
BPMH_DEV float2 prop_f32_beta(float kx2, float ky2, float a, float dz, float* beta) {
  const float tmp = f_sub(f_sub(a, kx2), ky2);
  const float h0 = tmp < 0.f ? 0.f : __fsqrt_rn(tmp);
  const float2 cs = sincos_f32_via_f64(f_mul(-dz, h0));
  *beta = tmp > 0.f ? h0 : 0.f;
  return tmp >= 0.f ? cs : make_float2(0.f, 0.f);
}

// Weighted mean (bpm_3d.py:149-159,302-303): dn0 = dn_sum/u_sum over the nx
// per-column partials of item b (x = dn part, y = u part) that column pass A
// wrote.  Column pass B forms it once per work-group: lane p of the first
// wavefront sums partials p, p+64, ... and an xor butterfly completes the
// sum (a fixed order, so every work-group gets the same bits); the result is
// broadcast through LDS, and the work-group of tile 0 stores it for the row
// pass.  The kernel boundary after pass A orders the partials' stores before
// these loads.
// slot[0] receives dn0, slot[1] the weight sum sum|U| (the adjoint needs it).
BPMH_DEV float weighted_dn0_group(const float2* __restrict__ wp, int nparts, float* slot) {
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    float sd = 0.f, su = 0.f;
    for (int p = lane; p < nparts; p += 64) {
      const float2 v = wp[p];
      sd += v.x;
      su += v.y;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      sd += __shfl_xor(sd, o, 64);
      su += __shfl_xor(su, o, 64);
    }
    if (lane == 0) {
      slot[0] = sd / su;
      slot[1] = su;
    }
  }
  __syncthreads();
  return slot[0];
}

// Sum of n1 floats at p1 and of n2 floats at p2 by the first wavefront in a
// fixed order (lane l takes elements l, l+64, ..., then an xor butterfly),
// so every work-group of an item gets the same bits; returned to all
// threads through slot[0..1].
BPMH_DEV float2 group_sum_pair(const float* __restrict__ p1, int n1, const float* __restrict__ p2,
                               int n2, float* slot) {
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    float s1 = 0.f, s2 = 0.f;
    for (int i = lane; i < n1; i += 64) s1 += p1[i];
    for (int i = lane; i < n2; i += 64) s2 += p2[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      s1 += __shfl_xor(s1, o, 64);
      s2 += __shfl_xor(s2, o, 64);
    }
    if (lane == 0) {
      slot[0] = s1;
      slot[1] = s2;
    }
  }
  __syncthreads();
  return make_float2(slot[0], slot[1]);
}

// Sum over the T threads of each column of a tile, in a fixed binary-tree
// order that depends only on T (not on how many columns share the
// work-group), so per-column partial sums are identical for every launch
// shape.  Uses the tile's LDS lines as scratch; returns the sum in t == 0.
template <int T>
BPMH_DEV float2 column_sum2(float2 v, float2* __restrict__ col_lds, int t) {
  __syncthreads();  // the FFT's last reads of col_lds are done
  col_lds[t] = v;
  __syncthreads();
#pragma unroll
  for (int s = T / 2; s > 0; s >>= 1) {
    if (t < s) col_lds[t] = c_add(col_lds[t], col_lds[t + s]);
    __syncthreads();
  }
  return col_lds[0];
}

// ----------------------------------------------------------------------------
// row pass
// ----------------------------------------------------------------------------
struct RowArgs {
  float2* plane;           // (rows = batch*ny, nx)
  const float2* tw;        // stage-major twiddles of the nx-point transform
  const void* dn;          // (nz, ny, nx) per batch item, kind dn_kind
  long long dn_bstride;    // elements between batch items' (or groups') dn (0: shared)
  // ragged batch: items [gend[g-1], gend[g]) read dn volume g; ngroups = 0:
  // item b reads volume b
  int ngroups;
  int gend[BPMH_MAX_GROUPS];
  int dn_kind;
  int ny, nz, z;
  float unit_k;            // float32(k0*dz)
  float inv_n;             // 1/(nx*ny)
  int absorb;
  int dn0_mode;            // BPMH_MEAN_*: 0 none, 1 per-plane array, 2 weighted
  const float* dn0_arr;    // simple mode: float32 dn0 per plane
  const float* dn0_item;   // weighted: dn0 of this step per item (column pass A)
  float2* out_field;       // (batch, nz, ny, nx) or null
  float* out_intens;       // (batch, nz, ny, nx) or null
  float* acc;              // (batch, ny, nx) or null: sum_z |u_z|^2
  int acc_init;
  int inverse;             // input is F_x u: inverse FFT first
  int screen;              // apply dn[z] screen
  int fwd_items;           // items b < fwd_items get the forward FFT_x before storing
  // adjoint step: v is the gradient G_z (torch convention, 2 dL/du*); the
  // row pass writes unit_k*Im(conj(G_z) u_z) per item and multiplies by the
  // conjugate screen
  int adjoint;
  const float2* useg;      // forward planes of the segment (batch, nseg, ny, nx)
  long long useg_bstride;  // elements between items in useg
  int useg_z;              // plane of z inside the segment
  float* gitem;            // (batch, ny, nx) per-item gradient of dn[z] (2 floats per voxel for complex dn)
};

// dn volume of item b: b itself, or in a ragged batch the group holding b
// (a few scalar compares; b is uniform over a work-group).
BPMH_DEV int dn_volume(const RowArgs& a, int b) {
  if (a.ngroups == 0) return b;
  int g = 0;
  while (g + 1 < a.ngroups && b >= a.gend[g]) ++g;
  return g;
}

template <int NX, int V, int EE>
__global__ void __launch_bounds__((RowCfg<NX, V, EE>::THREADS))
k_row(RowArgs a) {
  using Cfg = RowCfg<NX, V, EE>;
  constexpr int E = Cfg::E, T = Cfg::T, LINES = Cfg::LINES;
  // a line of at most 64 threads lies inside one wavefront (T divides 64)
  constexpr bool WAVE = T <= 64;
  __shared__ float2 lds[LINES][Cfg::ROW_LINE];
  const int line = threadIdx.x / T;
  const int t = threadIdx.x - line * T;
  const int row = blockIdx.x * LINES + line;  // in [0, batch*ny)
  const int b = row / a.ny;
  const int y = row - b * a.ny;
  float2* p = a.plane + (size_t)row * NX;
  const size_t plane_off = (size_t)y * NX;

  // Every global load of the pass is issued here, before the first barrier,
  // so their latencies overlap instead of adding up behind the transforms.
  float2 v[E];
#pragma unroll
  for (int m = 0; m < E; ++m) v[m] = ld_stream(p + t + m * T);
  Twiddles<NX, E> tw;
  const bool forward = b < a.fwd_items;  // uniform: a work-group's rows are one item's
  if (a.inverse || forward) tw.load(t, a.tw);
  float2 d[E];
  if (a.screen) {
    const size_t slice = ((size_t)a.z * a.ny + y) * NX + (size_t)dn_volume(a, b) * a.dn_bstride;
#pragma unroll
    for (int m = 0; m < E; ++m) d[m] = load_dn(a.dn, a.dn_kind, slice + t + m * T);
  }
  float accv[E];
  float* acc_row = a.acc ? a.acc + (size_t)b * a.ny * NX + plane_off : nullptr;
  if (acc_row && !a.acc_init) {
#pragma unroll
    for (int m = 0; m < E; ++m) accv[m] = acc_row[t + m * T];
  }
  float2 ua[E];
  if (a.adjoint) {
    const float2* u = a.useg + (size_t)b * a.useg_bstride + (size_t)a.useg_z * a.ny * NX + plane_off;
#pragma unroll
    for (int m = 0; m < E; ++m) ua[m] = u[t + m * T];
  }

  if (a.inverse) {
    line_fft<NX, E, true, SwzLayout<4>, WAVE>(v, lds[line], t, tw);
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = c_scale(v[m], a.inv_n);
  }

  if (a.screen) {
    float dn0 = 0.f;
    if (a.dn0_mode == BPMH_MEAN_SIMPLE) dn0 = a.dn0_arr[a.z];
    else if (a.dn0_mode == BPMH_MEAN_WEIGHTED) dn0 = a.dn0_item[b];
    const int disty = min(a.ny - y - 1, y);
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int x = t + m * T;
      if (a.dn_kind == BPMH_DN_C64) {
        // cexp((0,-k)*(dn - dn0)) = e^{k Im} (cos(-k Re), sin(-k Re)), no taper
        const float re = f_mul(a.unit_k, d[m].y);
        const float im = f_mul(-a.unit_k, f_sub(d[m].x, dn0));
        float s, c;
        sincos_fast(im, &s, &c);
        const float mag = expf(re);
        if (a.adjoint) {
          // u = s v with s = e^{k Im dn} e^{-i k Re dn}: dL/dRe dn = k Im(conj(G) u),
          // dL/dIm dn = k Re(conj(G) u), and G_v = conj(s) G
          const float2 w = make_float2(v[m].x * ua[m].x + v[m].y * ua[m].y,
                                       v[m].x * ua[m].y - v[m].y * ua[m].x);
          reinterpret_cast<float2*>(a.gitem)[((size_t)b * a.ny + y) * NX + x] =
              make_float2(a.unit_k * w.y, a.unit_k * w.x);
          v[m] = c_mulc(v[m], make_float2(mag * c, mag * s));
        } else {
          v[m] = c_mul(v[m], make_float2(mag * c, mag * s));
        }
      } else {
        const float ph = f_mul(-a.unit_k, f_sub(d[m].x, dn0));
        float s, c;
        sincos_fast(ph, &s, &c);
        float2 r;
        if (a.adjoint) {
          a.gitem[((size_t)b * a.ny + y) * NX + x] = a.unit_k * (v[m].x * ua[m].y - v[m].y * ua[m].x);
          r = c_mulc(v[m], make_float2(c, s));
        } else {
          r = c_mul(v[m], make_float2(c, s));
        }
        if (a.absorb > 0) {  // uniform: no taper work at all without one
          const int dist = min(min(NX - x - 1, x), disty);
          if (dist < a.absorb) {
            const float arg = f_div(f_mul(BPMH_PI_F, (float)dist), (float)a.absorb);
            const float tp = f_mul(0.5f, f_sub(1.f, (float)cos((double)arg)));
            r = make_float2(r.x * tp, r.y * tp);
          }
        }
        v[m] = r;
      }
    }
  }

  if (a.out_field) {
    float2* o = a.out_field + ((size_t)b * a.nz + a.z) * a.ny * NX + plane_off;
#pragma unroll
    for (int m = 0; m < E; ++m) st_stream(o + t + m * T, v[m]);
  }
  if (a.out_intens) {
    float* o = a.out_intens + ((size_t)b * a.nz + a.z) * a.ny * NX + plane_off;
#pragma unroll
    for (int m = 0; m < E; ++m) st_stream(o + t + m * T, v[m].x * v[m].x + v[m].y * v[m].y);
  }
  if (acc_row) {
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const float i2 = v[m].x * v[m].x + v[m].y * v[m].y;
      acc_row[t + m * T] = a.acc_init ? i2 : accv[m] + i2;
    }
  }

  if (forward) line_fft<NX, E, false, SwzLayout<4>, WAVE>(v, lds[line], t, tw);
#pragma unroll
  for (int m = 0; m < E; ++m) st_stream(p + t + m * T, v[m]);
}

// ----------------------------------------------------------------------------
// column pass
// ----------------------------------------------------------------------------
struct ColArgs {
  float2* plane;           // (batch, ny, nx)
  const float2* tw;        // stage-major twiddles of the ny-point transform
  int nx, nz, z;
  int hmode;               // 0: table H, 1: formula with n_eff[z], 2: formula with n0+weighted dn0
  const float2* H;         // (ny, nx)
  const float* n_eff;      // per plane, float32
  float n0, k0, dx, dy, dz;
  const void* dn;          // weighted: slice 0 of each item's dn is the weight
  long long dn_bstride;
  int dn_kind;
  float2* wpart;           // weighted per-column partials (batch, nx)
  float* dn0_item;         // weighted: dn0 per item, stored by pass B for the row pass
  const float* sw;         // scattering weights (ny, nx) or null
  const float* gw;
  const float2* pw;        // plane-wave DC per plane (nz)
  float2* spart;           // scattering per-column partials (batch, nz, nx)
  int conj_h;              // adjoint: multiply by conj(H)
  // forward recorded for the adjoint (null: nothing recorded): PHASE 1, and
  // PHASE 0 before H, also store U = F u of this step (item stride
  // kout_bstride); PHASE 2 stores this step's dn0 and sum|U| per item
  float2* kout;
  long long kout_bstride;
  float* rec_c;
  float* rec_w;
  // weighted adjoint (PHASE 3, 4) of the step whose forward k-space plane is
  // U (uk, item stride uk_bstride) with dn0 c = cz[b] and sum|U| = wz[b]
  const float2* uk;
  long long uk_bstride;
  const float* cz;
  const float* wz;
  float* s2part;           // (batch, nx): per column sum Im(conj(X) H U) / beta
  const float* s1part;     // (batch, ny): per row sum of the screen gradient
  float* gd_item;          // (batch, ny, nx): += dL/d(weight slice) per item
  // simple mean (PHASE 3, 4): dn0 = mean(dn[z]) does not depend on the
  // field: n = n_eff[z], no |U| term, and PHASE 4 stores dL/ddn0 per item
  int simple;
  float* cg_out;
};

// PHASE 0: FFT_y, H, IFFT_y.  PHASE 1: FFT_y, weighted partials, store.
// PHASE 2: weighted dn0, H(n0+dn0), IFFT_y.
// Adjoint of a weighted step (X = F G_v, the k-space gradient):
// PHASE 3: FFT_y, per-column partials of sum Im(conj(X) H U)/beta, store X.
// PHASE 4: dL/ddn0 from the partials, conj(H) X plus the term of dn0's
//          dependence on |U|, the weight-slice gradient, IFFT_y (simple mean:
//          conj(H) X, IFFT_y, dL/ddn0 stored).
// PHASE 5: PHASE 0 with H evaluated on the fly (hmode 1, the simple mean),
//          half the rows per thread as in PHASE 2.
template <int NY, int V, int PHASE, int EE>
__global__ void __launch_bounds__((ColCfg<NY, V, EE>::THREADS))
k_col(ColArgs a) {
  using Cfg = ColCfg<NY, V, EE>;
  constexpr int E = Cfg::E, T = Cfg::T, COLS = Cfg::COLS;
  constexpr bool FWD = PHASE == 0 || PHASE == 1 || PHASE == 3 || PHASE == 5;  // forward FFT_y
  constexpr bool INV = PHASE == 0 || PHASE == 2 || PHASE == 4 || PHASE == 5;  // H, inverse FFT_y
  constexpr bool ADJW = PHASE == 3 || PHASE == 4;
  // each thread holds P adjacent columns; the LC lane columns of a 16-lane
  // group are the LDS lines it touches together
  constexpr int P = Cfg::P, LC = Cfg::LC;
  constexpr bool SPLIT = Cfg::SPLIT;
  // swizzled lines 16/LC float2 apart, or padded lines 9 float2 longer
  // than padded: either way the LC lines a 16-lane group touches together
  // start in different LDS banks.  Two lane columns swizzle when each
  // thread holds two columns (1 % faster than padding at 1024^2 in batches)
  // and pad otherwise (the swizzle was slower there).
  constexpr bool SWZ = LC >= 4 || (P == 2 && LC >= 2);
  static_assert(COLS <= 16, "at most 16 columns per tile");
  constexpr int LINE = SWZ ? NY + 16 / LC : Cfg::COL_LINE;
  __shared__ float2 lds[COLS][LINE];
  using CL = std::conditional_t<SWZ, SwzLayout<4 - ilog2c(LC)>, PadLayout<Cfg::COL_SH>>;
  const int ntiles = a.nx / COLS;
  const int bx = blockIdx.x;
  // adjacent tiles share cache lines: keep them on one XCD (blocks are
  // dealt round-robin over the 8 XCDs)
  const int tile = (ntiles % 8 == 0) ? (bx & 7) * (ntiles >> 3) + (bx >> 3) : bx;
  const int b = blockIdx.y;
  const int c = threadIdx.x % LC;
  const int t = threadIdx.x / LC;
  const int x0 = tile * COLS + c * P;  // the thread's columns x0 .. x0+P-1
  float2* p = a.plane + (size_t)b * NY * a.nx + x0;

  // all global loads first (plane, twiddles, the propagator quadrant); with
  // P = 2 and 16-byte accesses the host guarantees a 16-byte aligned plane
  // (x0 and nx are even)
  float2 v[P][E];
  Twiddles<NY, E> tw;
  // split columns: the twiddles first, so waiting for the first column's
  // data (loads complete in issue order) does not wait for the second's
  if constexpr (SPLIT) tw.load(t, a.tw);
  if constexpr (SPLIT) {
#pragma unroll
    for (int q = 0; q < P; ++q)
#pragma unroll
      for (int m = 0; m < E; ++m) v[q][m] = p[(size_t)(t + m * T) * a.nx + q];
  } else {
#pragma unroll
    for (int m = 0; m < E; ++m) {
      float2* r = p + (size_t)(t + m * T) * a.nx;
      if constexpr (P == 2) {
        const float4 w = *reinterpret_cast<const float4*>(r);
        v[0][m] = make_float2(w.x, w.y);
        v[1][m] = make_float2(w.z, w.w);
      } else {
        v[0][m] = *r;
      }
    }
  }
  if constexpr (!SPLIT) tw.load(t, a.tw);
  // H depends on kx^2 and ky^2 only, so H[j][i] = H[(ny-j)%ny][(nx-i)%nx]
  // holds bit for bit; reading the quadrant j<=ny/2, i<=nx/2 keeps the
  // table's footprint at a quarter (L2-resident per XCD strip).
  // adjoint: the forward k-space plane U (and the weight slice) of the step
  float2 uv[P][E];
  float dv[P][E];
  if constexpr (ADJW) {
    const float2* uk = a.uk + (size_t)b * a.uk_bstride + x0;
#pragma unroll
    for (int q = 0; q < P; ++q)
#pragma unroll
      for (int m = 0; m < E; ++m) uv[q][m] = uk[(size_t)(t + m * T) * a.nx + q];
    if constexpr (PHASE == 4) {
      const size_t s0 = (size_t)b * a.dn_bstride + x0;
#pragma unroll
      for (int q = 0; q < P; ++q)
#pragma unroll
        for (int m = 0; m < E; ++m)
          dv[q][m] = a.dn ? load_dn(a.dn, a.dn_kind, s0 + (size_t)(t + m * T) * a.nx + q).x : 0.f;
    }
  }
  float2 hv[P][E];
  if (INV && a.hmode == 0) {
#pragma unroll
    for (int q = 0; q < P; ++q) {
      const int qx = min(x0 + q, a.nx - x0 - q);
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int ky = t + m * T;
        hv[q][m] = a.H[(size_t)min(ky, NY - ky) * a.nx + qx];
      }
    }
  }

  float dn0w = 0.f;
  if constexpr (PHASE == 2) {
    __shared__ float dn0_slot[2];
    dn0w = weighted_dn0_group(a.wpart + (size_t)b * a.nx, a.nx, dn0_slot);
    if (bx == 0 && threadIdx.x == 0) {
      a.dn0_item[b] = dn0w;
      if (a.rec_c) {
        a.rec_c[b] = dn0w;
        a.rec_w[b] = dn0_slot[1];
      }
    }
  }
  if constexpr (ADJW) dn0w = a.simple ? 0.f : a.cz[b];  // the forward's dn0 of this step
  // the step's refractive index: weighted n0 + dn0 (as PHASE 2), simple n_eff[z]
  float n_adj = 0.f;
  if constexpr (ADJW) n_adj = a.simple ? a.n_eff[a.z] : f_add(a.n0, dn0w);
  float cgrad = 0.f, inv_w = 0.f, n_pts = 0.f;
  if constexpr (PHASE == 4) {
    // dL/ddn0 = -sum_x (screen gradient) + dz (n0+dn0) k0^2 / N sum_k Im(conj(X) H U)/beta
    __shared__ float cg_slot[2];
    const float2 sums = group_sum_pair(a.s1part + (size_t)b * NY, NY, a.s2part + (size_t)b * a.nx,
                                       a.nx, cg_slot);
    n_pts = (float)a.nx * (float)NY;
    cgrad = -sums.x + a.dz * n_adj * a.k0 * a.k0 / n_pts * sums.y;
    if (a.simple) {
      if (bx == 0 && threadIdx.x == 0) a.cg_out[b] = cgrad;
    } else {
      inv_w = 1.f / a.wz[b];
    }
  }

  if constexpr (FWD) {
#pragma unroll
    for (int q = 0; q < P; ++q) line_fft<NY, E, false, CL>(v[q], lds[q * LC + c], t, tw);
  }
  if constexpr (PHASE == 0 || PHASE == 5) {
    // forward recorded for the adjoint: U = F u of this step, before H
    if (a.kout) {
      float2* ko = a.kout + (size_t)b * a.kout_bstride + x0;
#pragma unroll
      for (int q = 0; q < P; ++q)
#pragma unroll
        for (int m = 0; m < E; ++m) ko[(size_t)(t + m * T) * a.nx + q] = v[q][m];
    }
  }

#pragma unroll
  for (int q = 0; q < P; ++q) {
    const int x = x0 + q;
    float2* line = lds[q * LC + c];
    if constexpr (PHASE == 1) {
      const size_t s0 = (size_t)b * a.dn_bstride;
      float2 acc = make_float2(0.f, 0.f);
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int ky = t + m * T;
        const float au = hypotf(v[q][m].x, v[q][m].y);
        const float d = a.dn ? load_dn(a.dn, a.dn_kind, s0 + (size_t)ky * a.nx + x).x : 0.f;
        acc.x += au * d;
        acc.y += au;
      }
      acc = column_sum2<T>(acc, line, t);
      if (t == 0) a.wpart[(size_t)b * a.nx + x] = acc;
    } else if constexpr (PHASE == 3) {
      // dH/ddn0 = H * (-i dz (n0+dn0) k0^2 / beta): collect Im(conj(X) H U)/beta
      const float an = n_term_f32(n_adj, a.k0);
      const float kx2 = kcomp2_f32(x, a.nx, a.dx);
      // H and beta for half the rows, the mirrored rows (same values bit for
      // bit) from the partner thread through the line, as in the H passes
      float2 h[E];
      float bt[E];
#pragma unroll
      for (int m = 0; m < E / 2; ++m)
        h[m] = prop_f32_beta(kx2, kcomp2_f32(t + m * T, NY, a.dy), an, a.dz, &bt[m]);
      if (t == 0)
        h[E / 2] = prop_f32_beta(kx2, kcomp2_f32(NY / 2, NY, a.dy), an, a.dz, &bt[E / 2]);
      __syncthreads();  // earlier users of the line are done with it
#pragma unroll
      for (int m = 0; m < E / 2; ++m) {
        line[m * T + t] = h[m];
        line[NY / 2 + m * T + t] = make_float2(bt[m], 0.f);
      }
      __syncthreads();
#pragma unroll
      for (int m = E / 2; m < E; ++m) {
        if (t == 0) {
          if (m > E / 2) {
            h[m] = h[E - m];
            bt[m] = bt[E - m];
          }
        } else {
          h[m] = line[(E - 1 - m) * T + T - t];
          bt[m] = line[NY / 2 + (E - 1 - m) * T + T - t].x;
        }
      }
      float acc = 0.f;  // (column_sum2 waits for every reader of the line)
#pragma unroll
      for (int m = 0; m < E; ++m) {
        if (bt[m] > 0.f) {
          const float2 hu = c_mul(h[m], uv[q][m]);
          acc += (v[q][m].x * hu.y - v[q][m].y * hu.x) / bt[m];
        }
      }
      const float2 tot = column_sum2<T>(make_float2(acc, 0.f), line, t);
      if (t == 0) a.s2part[(size_t)b * a.nx + x] = tot.x;
    } else {
      float an = 0.f, kx2 = 0.f;
      if (a.hmode != 0) {
        float n = ADJW ? n_adj : (a.hmode == 1) ? a.n_eff[a.z] : f_add(a.n0, dn0w);
        an = n_term_f32(n, a.k0);
        kx2 = kcomp2_f32(x, a.nx, a.dx);
      }
      if constexpr (PHASE == 2 || PHASE == 4 || PHASE == 5) {
        // H(n0 + dn0) (weighted pass B, its adjoint) or H(n_eff[z]) (simple
        // mean, its adjoint) on the fly for every row.  H(ky) == H(NY - ky)
        // bit for bit (kcomp2_f32 of both rows is the same float), and the
        // rows of thread t mirror those of thread T - t (element m <->
        // E-1-m; thread 0 mirrors its own, m <-> E-m, and holds row NY/2
        // alone), so each thread evaluates its first E/2 rows (thread 0 also
        // row NY/2) and takes the others from its partner through the
        // column's LDS line (free before the inverse transform).
        float2 h[E];
#pragma unroll
        for (int m = 0; m < E / 2; ++m)
          h[m] = prop_f32(kx2, kcomp2_f32(t + m * T, NY, a.dy), an, a.dz);
        if (t == 0) h[E / 2] = prop_f32(kx2, kcomp2_f32(NY / 2, NY, a.dy), an, a.dz);
        __syncthreads();  // earlier users of the line are done with it
#pragma unroll
        for (int m = 0; m < E / 2; ++m) line[m * T + t] = h[m];
        __syncthreads();
        const int pt = T - t;
#pragma unroll
        for (int m = E / 2; m < E; ++m) {
          if (t == 0) {
            if (m > E / 2) h[m] = h[E - m];
          } else {
            h[m] = line[(E - 1 - m) * T + pt];
          }
        }
        __syncthreads();  // before the line is written again
#pragma unroll
        for (int m = 0; m < E; ++m)
          v[q][m] = a.conj_h ? c_mulc(v[q][m], h[m]) : c_mul(v[q][m], h[m]);
      } else {
#pragma unroll
        for (int m = 0; m < E; ++m) {
          const int ky = t + m * T;
          const float2 h =
              (a.hmode == 0) ? hv[q][m] : prop_f32(kx2, kcomp2_f32(ky, NY, a.dy), an, a.dz);
          v[q][m] = a.conj_h ? c_mulc(v[q][m], h) : c_mul(v[q][m], h);
        }
      }
      if (PHASE == 4 && !a.simple) {
        // dn0 = sum|U| d / sum|U| depends on U through |U|:
        // G_U += dL/ddn0 (d - dn0) / W * U/|U| (times N: X is N G_V), and
        // dL/dd += dL/ddn0 |U| / W
        float* gd = a.gd_item + (size_t)b * NY * a.nx + x;
#pragma unroll
        for (int m = 0; m < E; ++m) {
          const int ky = t + m * T;
          const float au = hypotf(uv[q][m].x, uv[q][m].y);
          if (au > 0.f) {
            const float coef = n_pts * cgrad * (dv[q][m] - dn0w) * inv_w / au;
            v[q][m].x += coef * uv[q][m].x;
            v[q][m].y += coef * uv[q][m].y;
          }
          gd[(size_t)ky * a.nx] += cgrad * au * inv_w;
        }
      }
      if (a.sw) {
        // scattering reductions of bpm_3d.py:243-247,335-341 on the k-space
        // plane after H: sum w |U - delta_{k,0} plain|^2
        float2 acc = make_float2(0.f, 0.f);
        const float2 pl = a.pw[a.z];
#pragma unroll
        for (int m = 0; m < E; ++m) {
          const int ky = t + m * T;
          float2 f = v[q][m];
          if (ky == 0 && x == 0) f = c_sub(f, pl);
          const float af = hypotf(f.x, f.y);
          const float a2 = af * af;
          const size_t k = (size_t)ky * a.nx + x;
          acc.x += a.sw[k] * a2;
          acc.y += a.gw[k] * a2;
        }
        acc = column_sum2<T>(acc, line, t);
        if (t == 0) a.spart[((size_t)b * a.nz + a.z) * a.nx + x] = acc;
        __syncthreads();  // the line is the inverse transform's exchange buffer
      }
    }
  }
  if constexpr (PHASE == 1) {
    // weighted forward recorded for the adjoint: U = F u of this step
    if (a.kout) {
      float2* ko = a.kout + (size_t)b * a.kout_bstride + x0;
#pragma unroll
      for (int q = 0; q < P; ++q)
#pragma unroll
        for (int m = 0; m < E; ++m) ko[(size_t)(t + m * T) * a.nx + q] = v[q][m];
    }
  }
  if constexpr (SPLIT) {
    // each column stored as soon as its inverse transform is done
#pragma unroll
    for (int q = 0; q < P; ++q) {
      if constexpr (INV) line_fft<NY, E, true, CL>(v[q], lds[q * LC + c], t, tw);
#pragma unroll
      for (int m = 0; m < E; ++m) p[(size_t)(t + m * T) * a.nx + q] = v[q][m];
    }
  } else {
    if constexpr (INV) {
#pragma unroll
      for (int q = 0; q < P; ++q) line_fft<NY, E, true, CL>(v[q], lds[q * LC + c], t, tw);
    }
#pragma unroll
    for (int m = 0; m < E; ++m) {
      float2* r = p + (size_t)(t + m * T) * a.nx;
      if constexpr (P == 2)
        *reinterpret_cast<float4*>(r) = make_float4(v[0][m].x, v[0][m].y, v[1][m].x, v[1][m].y);
      else
        *r = v[0][m];
    }
  }
}

// ----------------------------------------------------------------------------
// propagator tables and small helpers
// ----------------------------------------------------------------------------
// kind F32: compute_propagator (bpm_3d_kernels.cl:408-432), float32 formula.
// kind F64: host float64 propagator (bpm_3d.py:163-181, bpmclass/bpm3d.py:88-107)
//           rounded to complex64.
__global__ void k_prop_table(float2* H, int nx, int ny, int kind, double dx, double dy,
                             double dz, double n0, double lam) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= nx) return;
  const double k0 = 2. * M_PI / lam;
  float2 h;
  if (kind == BPMH_H_F32) {
    const float an = n_term_f32((float)n0, (float)k0);
    h = prop_f32(kcomp2_f32(i, nx, (float)dx), kcomp2_f32(j, ny, (float)dy), an, (float)dz);
  } else {
    // numpy: 2*pi*fftfreq(n,d) = 2*pi*(m * (1/(n*d)))
    const int mi = (i < (nx - 1) / 2 + 1) ? i : i - nx;
    const int mj = (j < (ny - 1) / 2 + 1) ? j : j - ny;
    const double kx = (2. * M_PI) * ((double)mi * (1.0 / (nx * dx)));
    const double ky = (2. * M_PI) * ((double)mj * (1.0 / (ny * dy)));
    const double tmp = (n0 * n0) * (k0 * k0) - kx * kx - ky * ky;
    if (tmp < 0.) {
      h = make_float2(0.f, 0.f);
    } else {
      double s, c;
      sincos(-dz * sqrt(tmp), &s, &c);
      h = make_float2((float)c, (float)s);
    }
  }
  H[(size_t)j * nx + i] = h;
}

__global__ void k_scatter_final(const float2* spart, float* out, int nz, int nparts) {
  const int z = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (z >= nz) return;
  float2 s = make_float2(0.f, 0.f);
  if (z > 0) {
    const float2* q = spart + ((size_t)b * nz + z) * nparts;
    for (int i = 0; i < nparts; ++i) s = c_add(s, q[i]);
  }
  out[((size_t)b * nz + z) * 2 + 0] = s.x;
  out[((size_t)b * nz + z) * 2 + 1] = s.y;
}

// ----------------------------------------------------------------------------
// plans
// ----------------------------------------------------------------------------
// Kernel kinds timed by the optional HIP-event instrumentation.
enum { TK_ROW = 0, TK_COL = 1, TK_COL_A = 2, TK_COL_B = 3, TK_N = 4 };

// One captured launch sequence of bpmh_propagate: its argument key, the
// executable graph, and an event recorded after its latest launch (an
// executable is destroyed only once that event has completed, so a replay
// still running on some stream is never destroyed under it).
struct GraphEntry {
  std::vector<unsigned char> key;
  hipGraphExec_t exec;
  hipEvent_t done;
  unsigned long long used;  // recency stamp for eviction
};

// Executables kept per plan (least recently used is evicted).
constexpr int kGraphCache = 4;

struct bpmh_plan_s {
  int nx, ny;
  // stage-major twiddles of both radix schedules (index ealt), on the host
  // and their device copies
  std::vector<float2> htw_x[2], htw_y[2];
  float2* tw_x[2];
  float2* tw_y[2];
  void* work;                        // step scratch (allocated on demand)
  size_t work_bytes;
  int device;
  // timing: events recorded around every launch of bpmh_propagate
  int timing;
  std::vector<hipEvent_t> ev;   // pairs (start, stop)
  std::vector<int> ev_kind;     // kind of each pair
  size_t ev_used;               // pairs recorded since the last read-out
  // graph mode: the launch sequence of bpmh_propagate is captured once per
  // distinct argument list and replayed with hipGraphLaunch
  int row_v, col_v;             // launch-shape variants (autotuned)
  int graph_mode;
  hipStream_t cap_stream;
  std::vector<GraphEntry> graphs;
  unsigned long long graph_clock;
};

// Wait for an entry's latest replay, then release it.
static int graph_release(GraphEntry& g) {
  hipError_t e1 = hipEventSynchronize(g.done);
  hipError_t e2 = hipGraphExecDestroy(g.exec);
  hipError_t e3 = hipEventDestroy(g.done);
  g.exec = nullptr;
  g.done = nullptr;
  hipError_t e = e1 != hipSuccess ? e1 : (e2 != hipSuccess ? e2 : e3);
  if (e != hipSuccess) return bpmh_set_error(BPMH_ERR_HIP, hipGetErrorString(e));
  return BPMH_OK;
}

static int graph_release_all(bpmh_plan p) {
  int rc = BPMH_OK;
  for (GraphEntry& g : p->graphs) {
    const int r = graph_release(g);
    if (r && !rc) rc = r;
  }
  p->graphs.clear();
  return rc;
}

static bool pow2_in_range(int n) { return n >= 16 && n <= 4096 && (n & (n - 1)) == 0; }

// Twiddles exp(-2 pi i m/N).  Angles are reduced to the first octant so the
// symmetric values (m = 0, N/8, N/4, ...) are exact; among the float32
// neighbours of (cos, sin) the pair with |w|^2 closest to 1 is kept, since
// a magnitude error of a twiddle is applied again on every step.
static void make_twiddles(int N, std::vector<float2>& tw) {
  tw.resize(N);
  for (int m = 0; m < N; ++m) {
    // exact octant reduction of the angle 2 pi m / N
    const long long M = 8LL * m;  // angle in units of 2*pi/(8N): octant = M / N
    const int oct = (int)(M / N);
    const double frac = (double)(M - (long long)oct * N) / N;  // in [0,1)
    double c, s;
    const double phi = frac * (M_PI / 4.0);
    const double cp = cos(phi), sp = sin(phi), cq = cos(M_PI / 4.0 - phi), sq = sin(M_PI / 4.0 - phi);
    switch (oct & 7) {  // cos/sin of oct*pi/4 + phi
      case 0: c = cp; s = sp; break;
      case 1: c = sq; s = cq; break;
      case 2: c = -sp; s = cp; break;
      case 3: c = -cq; s = sq; break;
      case 4: c = -cp; s = -sp; break;
      case 5: c = -sq; s = -cq; break;
      case 6: c = sp; s = -cp; break;
      default: c = cq; s = -sq; break;
    }
    if (frac == 0.0) {  // exact multiples of pi/4
      const double r = sqrt(0.5);
      const double cc[8] = {1, r, 0, -r, -1, -r, 0, r};
      const double ss[8] = {0, r, 1, r, 0, -r, -1, -r};
      c = cc[oct & 7];
      s = ss[oct & 7];
    }
    const float c0 = (float)c, s0 = (float)-s;
    float bc = c0, bs = s0;
    double best = 1e300;
    for (int dc = -1; dc <= 1; ++dc) {
      for (int ds = -1; ds <= 1; ++ds) {
        const float cf = dc == 0 ? c0 : nextafterf(c0, dc > 0 ? 2.f : -2.f);
        const float sf = ds == 0 ? s0 : nextafterf(s0, ds > 0 ? 2.f : -2.f);
        const double mag = fabs((double)cf * cf + (double)sf * sf - 1.0);
        const double ang = fabs(atan2((double)sf, (double)cf) - atan2(-s, c));
        const double score = mag + 0.25 * ang;
        if (score < best) { best = score; bc = cf; bs = sf; }
      }
    }
    tw[m] = make_float2(bc, bs);
  }
}

// Stage-major twiddles for the line FFT of N points with E elements per
// thread (fft.hpp: Twiddles::load), built from the natural-order table.
static void make_stage_twiddles(int N, int E, std::vector<float2>& out) {
  std::vector<float2> full;
  make_twiddles(N, full);
  int n = 0, e = 0;
  while ((1 << n) < N) ++n;
  while ((1 << e) < E) ++e;
  const int S = n / e, R0 = 1 << (n - S * e);
  int ns = R0 > 1 ? R0 : E;
  const int nst = R0 > 1 ? S : S - 1;
  out.clear();
  for (int s = 0; s < nst; ++s) {
    for (int r = 1; r < E; ++r)
      for (int k = 0; k < ns; ++k) out.push_back(full[(size_t)r * k * (N / (ns * E))]);
    ns *= E;
  }
  if (out.empty()) out.push_back(make_float2(1.f, 0.f));
}

static int plan_upload(bpmh_plan p);

extern "C" int bpmh_plan_create(int nx, int ny, bpmh_plan* out) {
  if (!out) return bpmh_set_error(BPMH_ERR_INVALID, "null plan pointer");
  *out = nullptr;
  if (!pow2_in_range(nx) || !pow2_in_range(ny))
    return bpmh_set_error(BPMH_ERR_UNSUPPORTED,
                          "grid sizes must be powers of two in [16, 4096]");
  bpmh_plan p = new bpmh_plan_s();
  p->nx = nx;
  p->ny = ny;
  p->work = nullptr;
  p->work_bytes = 0;
  p->timing = 0;
  p->ev_used = 0;
  p->row_v = 1;
  p->col_v = 1;
  p->graph_mode = 0;
  p->cap_stream = nullptr;
  p->graph_clock = 0;
  p->device = -1;
  for (int e = 0; e < 2; ++e) {
    p->tw_x[e] = p->tw_y[e] = nullptr;
    make_stage_twiddles(nx, row_e_of(nx, e), p->htw_x[e]);
    make_stage_twiddles(ny, col_e_of(ny, e), p->htw_y[e]);
  }
  // device copies now when a device is present (on a machine without one
  // the plan stays host-only: its argument checks still run)
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) == hipSuccess && ndev > 0) {
    const int rc = plan_upload(p);
    if (rc) {
      delete p;
      return rc;
    }
  } else {
    (void)hipGetLastError();
  }
  *out = p;
  return BPMH_OK;
}

// Device copies of the twiddle tables on the current device, which becomes
// the plan's device (made at plan creation, or at the first propagation of a
// plan created without a device).
static int plan_upload(bpmh_plan p) {
  if (p->tw_x[0]) return BPMH_OK;
  HIP_TRY(hipGetDevice(&p->device));
  auto release = [p] {
    for (int e = 0; e < 2; ++e) {
      (void)hipFree(p->tw_x[e]);
      (void)hipFree(p->tw_y[e]);
      p->tw_x[e] = p->tw_y[e] = nullptr;
    }
  };
  for (int e = 0; e < 2; ++e) {
    if (hipMalloc(&p->tw_x[e], p->htw_x[e].size() * sizeof(float2)) != hipSuccess ||
        hipMalloc(&p->tw_y[e], p->htw_y[e].size() * sizeof(float2)) != hipSuccess) {
      release();
      return bpmh_set_error(BPMH_ERR_NOMEM, "hipMalloc failed for twiddles");
    }
    if (hipMemcpy(p->tw_x[e], p->htw_x[e].data(), p->htw_x[e].size() * sizeof(float2),
                  hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->tw_y[e], p->htw_y[e].data(), p->htw_y[e].size() * sizeof(float2),
                  hipMemcpyHostToDevice) != hipSuccess) {
      release();
      return bpmh_set_error(BPMH_ERR_HIP, "hipMemcpy failed for twiddles");
    }
  }
  return BPMH_OK;
}

extern "C" int bpmh_plan_destroy(bpmh_plan p) {
  if (!p) return BPMH_OK;
  for (int e = 0; e < 2; ++e) {
    (void)hipFree(p->tw_x[e]);
    (void)hipFree(p->tw_y[e]);
  }
  if (p->work) (void)hipFree(p->work);
  for (hipEvent_t e : p->ev) (void)hipEventDestroy(e);
  (void)graph_release_all(p);
  if (p->cap_stream) (void)hipStreamDestroy(p->cap_stream);
  delete p;
  return BPMH_OK;
}

extern "C" int bpmh_plan_set_graph(bpmh_plan p, int enable) {
  if (!p) return bpmh_set_error(BPMH_ERR_INVALID, "null plan");
  p->graph_mode = enable != 0;
  if (!p->graph_mode) return graph_release_all(p);
  return BPMH_OK;
}

extern "C" int bpmh_plan_set_timing(bpmh_plan p, int enable) {
  if (!p) return bpmh_set_error(BPMH_ERR_INVALID, "null plan");
  p->timing = enable != 0;
  p->ev_used = 0;
  return BPMH_OK;
}

extern "C" int bpmh_plan_kernel_times(bpmh_plan p, float* ms, int* counts) {
  if (!p || !ms || !counts) return bpmh_set_error(BPMH_ERR_INVALID, "null argument");
  for (int k = 0; k < TK_N; ++k) {
    ms[k] = 0.f;
    counts[k] = 0;
  }
  for (size_t i = 0; i < p->ev_used; ++i) {
    float t = 0.f;
    HIP_TRY(hipEventSynchronize(p->ev[2 * i + 1]));
    HIP_TRY(hipEventElapsedTime(&t, p->ev[2 * i], p->ev[2 * i + 1]));
    ms[p->ev_kind[i]] += t;
    counts[p->ev_kind[i]] += 1;
  }
  p->ev_used = 0;
  return BPMH_OK;
}

// Reserve event pairs for n more timed launches (outside any stream capture).
static int timing_reserve(bpmh_plan p, size_t n) {
  const size_t need = 2 * (p->ev_used + n);
  while (p->ev.size() < need) {
    hipEvent_t e;
    HIP_TRY(hipEventCreate(&e));
    p->ev.push_back(e);
  }
  if (p->ev_kind.size() < need / 2) p->ev_kind.resize(need / 2);
  return BPMH_OK;
}

// Plan scratch of at least `bytes` (grown, never shrunk).
static int plan_scratch(bpmh_plan p, size_t bytes, void** ptr) {
  if (bytes > p->work_bytes) {
    if (p->work) {
      HIP_TRY(hipDeviceSynchronize());
      HIP_TRY(hipFree(p->work));
      p->work = nullptr;
      p->work_bytes = 0;
    }
    bytes = (bytes + 255) & ~(size_t)255;
    if (hipMalloc(&p->work, bytes) != hipSuccess)
      return bpmh_set_error(BPMH_ERR_NOMEM, "hipMalloc failed for plan scratch");
    p->work_bytes = bytes;
  }
  *ptr = p->work;
  return BPMH_OK;
}

extern "C" int bpmh_propagator_table(int nx, int ny, int kind, double dx, double dy, double dz,
                                     double n0, double lam, float* H_, void* stream) {
  float2* H = (float2*)H_;
  if (!H) return bpmh_set_error(BPMH_ERR_INVALID, "null H");
  if (nx <= 0 || ny <= 0) return bpmh_set_error(BPMH_ERR_INVALID, "bad grid");
  if (kind != BPMH_H_F32 && kind != BPMH_H_F64)
    return bpmh_set_error(BPMH_ERR_INVALID, "unknown propagator kind");
  dim3 grid((nx + 255) / 256, ny);
  hipLaunchKernelGGL(k_prop_table, grid, dim3(256), 0, (hipStream_t)stream, H, nx, ny, kind, dx,
                     dy, dz, n0, lam);
  HIP_TRY(hipGetLastError());
  return BPMH_OK;
}

// ----------------------------------------------------------------------------
// launchers
// ----------------------------------------------------------------------------
typedef void (*row_fn)(const RowArgs&, int batch, hipStream_t);
typedef void (*col_fn)(const ColArgs&, int batch, hipStream_t);

template <int N, int V, int E>
static void launch_row(const RowArgs& a, int batch, hipStream_t s) {
  const int rows = batch * a.ny;
  hipLaunchKernelGGL((k_row<N, V, E>), dim3(rows / RowCfg<N, V, E>::LINES),
                     dim3(RowCfg<N, V, E>::THREADS), 0, s, a);
}
template <int N, int V, int PHASE, int E>
static void launch_col(const ColArgs& a, int batch, hipStream_t s) {
  hipLaunchKernelGGL((k_col<N, V, PHASE, E>), dim3(a.nx / ColCfg<N, V, E>::COLS, batch),
                     dim3(ColCfg<N, V, E>::THREADS), 0, s, a);
}

#define BPMH_SIZES(X) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096)

// Row launcher of variant v = V + 3*ealt for rows of N points.
template <int N>
static row_fn pick_row_n(int v) {
  constexpr int E0 = row_e_of(N, 0), E1 = row_e_of(N, 1);
  switch (v) {
    case 0: return launch_row<N, 0, E0>;
    case 2: return launch_row<N, 2, E0>;
    case 3: return launch_row<N, 0, E1>;
    case 4: return launch_row<N, 1, E1>;
    case 5: return launch_row<N, 2, E1>;
    default: return launch_row<N, 1, E0>;
  }
}

static row_fn pick_row(int n, int v) {
#define ROW_CASE(N) \
  case N:           \
    return pick_row_n<N>(v);
  switch (n) { BPMH_SIZES(ROW_CASE) }
#undef ROW_CASE
  return nullptr;
}

// Column launcher of variant v = V + 5*ealt for columns of N points.
template <int N, int PHASE>
static col_fn pick_col_n(int v) {
  constexpr int E0 = col_e_of(N, 0), E1 = col_e_of(N, 1);
  switch (v) {
    case 0: return launch_col<N, 0, PHASE, E0>;
    case 2: return launch_col<N, 2, PHASE, E0>;
    case 3: return launch_col<N, 3, PHASE, E0>;
    case 4: return launch_col<N, 4, PHASE, E0>;
    case 5: return launch_col<N, 0, PHASE, E1>;
    case 6: return launch_col<N, 1, PHASE, E1>;
    case 7: return launch_col<N, 2, PHASE, E1>;
    case 8: return launch_col<N, 3, PHASE, E1>;
    case 9: return launch_col<N, 4, PHASE, E1>;
    default: return launch_col<N, 1, PHASE, E0>;
  }
}

// Column phases of the weighted adjoint: one launch shape (V = 1) in the
// radix schedule of column variant v (its twiddle table is the plan's).
template <int PHASE>
static col_fn pick_col_adj(int n, int v) {
#define COLA_CASE(N)                                                                   \
  case N:                                                                              \
    return v >= kColShapes ? launch_col<N, 1, PHASE, col_e_of(N, 1)>                  \
                           : launch_col<N, 1, PHASE, col_e_of(N, 0)>;
  switch (n) { BPMH_SIZES(COLA_CASE) }
#undef COLA_CASE
  return nullptr;
}

template <int PHASE>
static col_fn pick_col(int n, int v) {
#define COL_CASE(N) \
  case N:           \
    return pick_col_n<N, PHASE>(v);
  switch (n) { BPMH_SIZES(COL_CASE) }
#undef COL_CASE
  return nullptr;
}

template <int N>
static int col_cols_n(int v) {
  constexpr int E0 = col_e_of(N, 0), E1 = col_e_of(N, 1);
  switch (v) {
    case 0: return ColCfg<N, 0, E0>::COLS;
    case 2: return ColCfg<N, 2, E0>::COLS;
    case 3: return ColCfg<N, 3, E0>::COLS;
    case 4: return ColCfg<N, 4, E0>::COLS;
    case 5: return ColCfg<N, 0, E1>::COLS;
    case 6: return ColCfg<N, 1, E1>::COLS;
    case 7: return ColCfg<N, 2, E1>::COLS;
    case 8: return ColCfg<N, 3, E1>::COLS;
    case 9: return ColCfg<N, 4, E1>::COLS;
    default: return ColCfg<N, 1, E0>::COLS;
  }
}

template <int N>
static int row_lines_n(int v) {
  constexpr int E0 = row_e_of(N, 0), E1 = row_e_of(N, 1);
  switch (v) {
    case 0: return RowCfg<N, 0, E0>::LINES;
    case 2: return RowCfg<N, 2, E0>::LINES;
    case 3: return RowCfg<N, 0, E1>::LINES;
    case 4: return RowCfg<N, 1, E1>::LINES;
    case 5: return RowCfg<N, 2, E1>::LINES;
    default: return RowCfg<N, 1, E0>::LINES;
  }
}

static int col_tiles(int nx, int ny, int v) {
#define TILE_CASE(N) \
  case N:            \
    return nx / col_cols_n<N>(v);
  switch (ny) { BPMH_SIZES(TILE_CASE) }
#undef TILE_CASE
  return 0;
}

// Column variants with V = 3 load and store 16 bytes per column pair: a
// plane that is only 8-byte aligned (a view at an odd element offset) runs
// V = 0 of the same radix schedule (one column per thread), which gives the
// same bits.
static int col_variant_for(int cv, const float2* plane) {
  return (cv % kColShapes == 3 && (reinterpret_cast<uintptr_t>(plane) & 15)) ? cv - 3 : cv;
}

static void variant_shape(int nx, int ny, int rv, int cv, int* lines, int* cols, int* row_e,
                          int* col_e) {
#define SHAPE_CASE(N)                 \
  case N:                             \
    *lines = row_lines_n<N>(rv);      \
    *row_e = row_e_of(N, rv / kRowShapes); \
    break;
  switch (nx) { BPMH_SIZES(SHAPE_CASE) }
#undef SHAPE_CASE
#define CE_CASE(N)                    \
  case N:                             \
    *col_e = col_e_of(N, cv / kColShapes); \
    break;
  switch (ny) { BPMH_SIZES(CE_CASE) }
#undef CE_CASE
  *cols = nx / col_tiles(nx, ny, cv);
}

extern "C" int bpmh_plan_set_variant(bpmh_plan p, int row_variant, int col_variant) {
  if (!p) return bpmh_set_error(BPMH_ERR_INVALID, "null plan");
  if (row_variant < 0 || row_variant >= 2 * kRowShapes || col_variant < 0 ||
      col_variant >= 2 * kColShapes)
    return bpmh_set_error(BPMH_ERR_INVALID, "row variants are 0 to 5; column variants 0 to 9");
  p->row_v = row_variant;
  p->col_v = col_variant;
  return BPMH_OK;
}

extern "C" int bpmh_plan_get_variant(bpmh_plan p, int* row_variant, int* col_variant,
                                     int* lines_per_group, int* cols_per_group, int* row_elems,
                                     int* col_elems) {
  if (!p || !row_variant || !col_variant || !lines_per_group || !cols_per_group || !row_elems ||
      !col_elems)
    return bpmh_set_error(BPMH_ERR_INVALID, "null argument");
  *row_variant = p->row_v;
  *col_variant = p->col_v;
  variant_shape(p->nx, p->ny, p->row_v, p->col_v, lines_per_group, cols_per_group, row_elems,
                col_elems);
  return BPMH_OK;
}

// Validated arguments of one bpmh_propagate call.
struct PropCall {
  int batch, nz;
  float2* plane;
  const void* dn;
  int dn_kind;
  long long dn_bstride, dn_numel;
  int mean_mode;
  const float2* H;
  const float* n_eff;
  const float* dn0;
  float n0, k0, dx, dy, dz, unit_k;
  int absorb;
  float2* out_field;
  float* out_intens;
  long long out_numel;
  float* out_sum;
  const float* scat_w;
  const float* g_w;
  const float2* plain;
  float* out_scat;
};

static int prop_validate(bpmh_plan plan, const PropCall& c) {
  char b[256];
  if (!plan) return bpmh_set_error(BPMH_ERR_INVALID, "null plan");
  if (!c.plane) return bpmh_set_error(BPMH_ERR_INVALID, "null plane");
  if (c.batch < 1 || c.nz < 1) return bpmh_set_error(BPMH_ERR_INVALID, "batch and nz must be >= 1");
  if (c.dn_kind < BPMH_DN_NONE || c.dn_kind > BPMH_DN_C64)
    return bpmh_set_error(BPMH_ERR_INVALID, "unknown dn kind");
  if (c.dn_kind != BPMH_DN_NONE && !c.dn) return bpmh_set_error(BPMH_ERR_INVALID, "null dn");
  if (c.dn_bstride < 0) return bpmh_set_error(BPMH_ERR_INVALID, "dn batch stride must be >= 0");
  if (c.mean_mode < BPMH_MEAN_IGNORE || c.mean_mode > BPMH_MEAN_WEIGHTED)
    return bpmh_set_error(BPMH_ERR_INVALID, "unknown mean mode");
  if (c.mean_mode == BPMH_MEAN_IGNORE && !c.H)
    return bpmh_set_error(BPMH_ERR_INVALID, "mean_mode ignore needs a propagator table");
  if (c.mean_mode == BPMH_MEAN_SIMPLE && (!c.n_eff || !c.dn0))
    return bpmh_set_error(BPMH_ERR_INVALID, "mean_mode simple needs n_eff and dn0 arrays");
  if (c.absorb < 0) return bpmh_set_error(BPMH_ERR_INVALID, "absorbing width must be >= 0");
  if (c.out_scat && (!c.scat_w || !c.g_w || !c.plain))
    return bpmh_set_error(BPMH_ERR_INVALID, "scattering output needs weights and plane-wave array");
  // the column grid's y extent is the batch (at most 65535)
  if (c.batch > 65535 || (long long)c.batch * plan->ny > (1LL << 30))
    return bpmh_set_error(BPMH_ERR_INVALID, "batch too large");
  const long long n = (long long)plan->nx * plan->ny;
  if (c.dn_kind != BPMH_DN_NONE) {
    // planes 0..nz-1 of every item (slice 0 also feeds the weighted mean)
    const long long need = (long long)(c.batch - 1) * c.dn_bstride + (long long)c.nz * n;
    if (c.dn_numel < need) {
      snprintf(b, sizeof b, "dn holds %lld elements; %d planes of %d item(s) at stride %lld read %lld",
               c.dn_numel, c.nz, c.batch, c.dn_bstride, need);
      return bpmh_set_error(BPMH_ERR_INVALID, b);
    }
  }
  if (c.out_field || c.out_intens) {
    const long long need = (long long)c.batch * c.nz * n;
    if (c.out_numel < need) {
      snprintf(b, sizeof b, "output volume holds %lld elements; %d item(s) x %d planes write %lld",
               c.out_numel, c.batch, c.nz, need);
      return bpmh_set_error(BPMH_ERR_INVALID, b);
    }
  }
  return BPMH_OK;
}

// Scratch of one call: weighted dn0 per item (batch floats), weighted
// per-column partials (batch*nx float2) and scattering per-column partials
// (batch*nz*nx float2, only with out_scat).
struct StepScratch {
  float2* wpart;
  float2* spart;
  float* dn0;
};

// 0 when the call needs none (no weighted mean, no scattering sums).
static size_t prop_scratch_bytes(bpmh_plan plan, const PropCall& c) {
  if (c.mean_mode != BPMH_MEAN_WEIGHTED && !c.out_scat) return 0;
  const size_t n_w = (size_t)c.batch * plan->nx;
  const size_t n_s = c.out_scat ? (size_t)c.batch * c.nz * plan->nx : 0;
  return (n_w + n_s) * sizeof(float2) + (((size_t)c.batch * 4 + 255) & ~(size_t)255);
}

static StepScratch prop_scratch_carve(bpmh_plan plan, const PropCall& c, void* base) {
  StepScratch r;
  const size_t n_w = (size_t)c.batch * plan->nx;
  const size_t n_s = c.out_scat ? (size_t)c.batch * c.nz * plan->nx : 0;
  const size_t seg = ((size_t)c.batch * 4 + 255) & ~(size_t)255;
  char* p = (char*)base;
  r.dn0 = (float*)p;
  r.wpart = (float2*)(p + seg);
  r.spart = r.wpart + n_w;
  (void)n_s;
  return r;
}

// Enqueue (or capture and replay) the whole z-loop of one call on stream s.
// may_allocate false: plain launches only (no graph capture or timing
// events, which create runtime objects).
static int prop_run(bpmh_plan plan, const PropCall& c, const StepScratch& sc, hipStream_t s,
                    bool may_allocate) {
  const int nx = plan->nx, ny = plan->ny;
  const int batch = c.batch, nz = c.nz;
  const row_fn row = pick_row(nx, plan->row_v);
  const int cv = col_variant_for(plan->col_v, c.plane);
  // the simple mean's propagator is evaluated in the column pass (PHASE 5)
  const col_fn col0 =
      c.mean_mode == BPMH_MEAN_SIMPLE ? pick_col<5>(ny, cv) : pick_col<0>(ny, cv);
  const col_fn colA = pick_col<1>(ny, cv);
  const col_fn colB = pick_col<2>(ny, cv);

  RowArgs ra;
  memset(&ra, 0, sizeof ra);
  ra.plane = c.plane;
  ra.tw = plan->tw_x[plan->row_v / kRowShapes];
  ra.dn = c.dn;
  ra.dn_bstride = c.dn_bstride;
  ra.dn_kind = c.dn_kind;
  ra.ny = ny;
  ra.nz = nz;
  ra.unit_k = c.unit_k;
  ra.inv_n = 1.0f / ((float)nx * (float)ny);
  ra.absorb = c.absorb;
  ra.dn0_mode = (c.dn_kind == BPMH_DN_NONE) ? 0 : c.mean_mode;
  ra.dn0_arr = c.dn0;
  ra.dn0_item = sc.dn0;
  ra.out_field = c.out_field;
  ra.out_intens = c.out_intens;
  ra.acc = c.out_sum;

  ColArgs ca;
  memset(&ca, 0, sizeof ca);
  ca.plane = c.plane;
  ca.tw = plan->tw_y[cv / kColShapes];
  ca.nx = nx;
  ca.nz = nz;
  ca.hmode = c.mean_mode;  // 0 table, 1 n_eff array, 2 weighted
  ca.H = c.H;
  ca.n_eff = c.n_eff;
  ca.n0 = c.n0;
  ca.k0 = c.k0;
  ca.dx = c.dx;
  ca.dy = c.dy;
  ca.dz = c.dz;
  ca.dn = c.dn;
  ca.dn_bstride = c.dn_bstride;
  ca.dn_kind = c.dn_kind;
  ca.wpart = sc.wpart;
  ca.dn0_item = sc.dn0;
  ca.sw = c.out_scat ? c.scat_w : nullptr;
  ca.gw = c.g_w;
  ca.pw = c.plain;
  ca.spart = sc.spart;

  const bool timing = may_allocate && plan->timing;
  if (timing) {
    const int rc = timing_reserve(plan, (size_t)nz * 3 + 1);
    if (rc) return rc;
  }
  const bool use_graph = may_allocate && plan->graph_mode && !plan->timing;
  const hipStream_t user = s;
  std::vector<unsigned char> key;
  if (use_graph) {
    auto put = [&key](const void* v, size_t n) {
      const unsigned char* q = (const unsigned char*)v;
      key.insert(key.end(), q, q + n);
    };
    put(&ra, sizeof ra);
    put(&ca, sizeof ca);
    put(&batch, sizeof batch);
    put(&nz, sizeof nz);
    put(&c.mean_mode, sizeof c.mean_mode);
    put(&c.out_scat, sizeof c.out_scat);
    put(&plan->row_v, sizeof plan->row_v);
    put(&plan->col_v, sizeof plan->col_v);
    for (GraphEntry& g : plan->graphs) {
      if (g.key == key) {
        HIP_TRY(hipGraphLaunch(g.exec, user));
        HIP_TRY(hipEventRecord(g.done, user));
        g.used = ++plan->graph_clock;
        return BPMH_OK;
      }
    }
    if (!plan->cap_stream) HIP_TRY(hipStreamCreateWithFlags(&plan->cap_stream, hipStreamNonBlocking));
    HIP_TRY(hipStreamBeginCapture(plan->cap_stream, hipStreamCaptureModeThreadLocal));
    s = plan->cap_stream;
  }
  auto timed = [&](int kind, auto&& launch) {
    if (!timing) {
      launch();
      return;
    }
    const size_t i = plan->ev_used++;
    plan->ev_kind[i] = kind;
    (void)hipEventRecord(plan->ev[2 * i], s);
    launch();
    (void)hipEventRecord(plan->ev[2 * i + 1], s);
  };

  // plane 0: outputs of u0, then FFT_x
  ra.z = 0;
  ra.acc_init = 1;
  ra.inverse = 0;
  ra.screen = 0;
  ra.fwd_items = nz > 1 ? batch : 0;
  timed(TK_ROW, [&] { row(ra, batch, s); });
  ra.acc_init = 0;
  ra.inverse = 1;
  ra.screen = c.dn_kind != BPMH_DN_NONE;
  for (int z = 1; z < nz; ++z) {
    ca.z = z;
    if (c.mean_mode == BPMH_MEAN_WEIGHTED) {
      timed(TK_COL_A, [&] { colA(ca, batch, s); });
      timed(TK_COL_B, [&] { colB(ca, batch, s); });
    } else {
      timed(TK_COL, [&] { col0(ca, batch, s); });
    }
    ra.z = z;
    ra.fwd_items = z < nz - 1 ? batch : 0;
    timed(TK_ROW, [&] { row(ra, batch, s); });
  }
  if (c.out_scat) {
    hipLaunchKernelGGL(k_scatter_final, dim3((nz + 63) / 64, batch), dim3(64), 0, s, sc.spart,
                       c.out_scat, nz, nx);
  }
  if (use_graph) {
    hipGraph_t g = nullptr;
    hipError_t le = hipGetLastError();
    hipError_t ce = hipStreamEndCapture(plan->cap_stream, &g);
    if (le != hipSuccess || ce != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return bpmh_set_error(BPMH_ERR_HIP, hipGetErrorString(le != hipSuccess ? le : ce));
    }
    GraphEntry e;
    e.key = std::move(key);
    e.exec = nullptr;
    e.done = nullptr;
    hipError_t ie = hipGraphInstantiate(&e.exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) return bpmh_set_error(BPMH_ERR_HIP, hipGetErrorString(ie));
    hipError_t ee = hipEventCreateWithFlags(&e.done, hipEventDisableTiming);
    if (ee != hipSuccess) {
      (void)hipGraphExecDestroy(e.exec);
      return bpmh_set_error(BPMH_ERR_HIP, hipGetErrorString(ee));
    }
    if ((int)plan->graphs.size() >= kGraphCache) {
      size_t lru = 0;
      for (size_t i = 1; i < plan->graphs.size(); ++i)
        if (plan->graphs[i].used < plan->graphs[lru].used) lru = i;
      const int rc = graph_release(plan->graphs[lru]);
      plan->graphs.erase(plan->graphs.begin() + lru);
      if (rc) {
        (void)hipGraphExecDestroy(e.exec);
        (void)hipEventDestroy(e.done);
        return rc;
      }
    }
    e.used = ++plan->graph_clock;
    plan->graphs.push_back(std::move(e));
    GraphEntry& ne = plan->graphs.back();
    HIP_TRY(hipGraphLaunch(ne.exec, user));
    HIP_TRY(hipEventRecord(ne.done, user));
    return BPMH_OK;
  }
  HIP_TRY(hipGetLastError());
  return BPMH_OK;
}

// ----------------------------------------------------------------------------
// memory floor of a step: the two passes' global loads and stores in their
// launch shapes, without the transforms
// ----------------------------------------------------------------------------
// Row pattern: each line of the row pass read and written with streaming
// accesses and its dn row read (20 B per voxel), one multiply in between.
template <int NX, int V, int EE>
__global__ void __launch_bounds__((RowCfg<NX, V, EE>::THREADS))
k_floor_row(float2* __restrict__ plane, const float* __restrict__ dn, int ny, int z) {
  using Cfg = RowCfg<NX, V, EE>;
  constexpr int E = Cfg::E, T = Cfg::T, LINES = Cfg::LINES;
  const int line = threadIdx.x / T;
  const int t = threadIdx.x - line * T;
  const int row = blockIdx.x * LINES + line;
  const int y = row % ny;
  float2* p = plane + (size_t)row * NX;
  const float* d = dn + ((size_t)z * ny + y) * NX;
  float2 v[E];
  float dv[E];
#pragma unroll
  for (int m = 0; m < E; ++m) v[m] = ld_stream(p + t + m * T);
#pragma unroll
  for (int m = 0; m < E; ++m) dv[m] = d[t + m * T];
#pragma unroll
  for (int m = 0; m < E; ++m) st_stream(p + t + m * T, c_scale(v[m], 1.f + dv[m]));
}

// Column pattern: the column pass's tiles (XCD-aware tile order, P columns
// per thread with 16-byte accesses when P = 2) read and written, 16 B per
// voxel.
template <int NY, int V, int EE>
__global__ void __launch_bounds__((ColCfg<NY, V, EE>::THREADS)) k_floor_col(float2* __restrict__ plane,
                                                                          int nx) {
  using Cfg = ColCfg<NY, V, EE>;
  constexpr int E = Cfg::E, T = Cfg::T, COLS = Cfg::COLS, P = Cfg::P, LC = Cfg::LC;
  const int ntiles = nx / COLS;
  const int bx = blockIdx.x;
  const int tile = (ntiles % 8 == 0) ? (bx & 7) * (ntiles >> 3) + (bx >> 3) : bx;
  const int c = threadIdx.x % LC;
  const int t = threadIdx.x / LC;
  float2* p = plane + (size_t)blockIdx.y * NY * nx + tile * COLS + c * P;
  constexpr bool WIDE = P == 2 && !Cfg::SPLIT;  // 16-byte accesses, as the pass
  float2 v[P][E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    float2* r = p + (size_t)(t + m * T) * nx;
    if constexpr (WIDE) {
      const float4 w = *reinterpret_cast<const float4*>(r);
      v[0][m] = make_float2(w.x, w.y);
      v[1][m] = make_float2(w.z, w.w);
    } else {
#pragma unroll
      for (int q = 0; q < P; ++q) v[q][m] = r[q];
    }
  }
#pragma unroll
  for (int m = 0; m < E; ++m) {
    float2* r = p + (size_t)(t + m * T) * nx;
    if constexpr (WIDE) {
      *reinterpret_cast<float4*>(r) = make_float4(-v[0][m].x, v[0][m].y, -v[1][m].x, v[1][m].y);
    } else {
#pragma unroll
      for (int q = 0; q < P; ++q) r[q] = make_float2(-v[q][m].x, v[q][m].y);
    }
  }
}

template <int N, int V, int E>
static void launch_floor_row(float2* plane, const float* dn, int batch, int ny, int z,
                             hipStream_t s) {
  hipLaunchKernelGGL((k_floor_row<N, V, E>), dim3(batch * ny / RowCfg<N, V, E>::LINES),
                     dim3(RowCfg<N, V, E>::THREADS), 0, s, plane, dn, ny, z);
}

template <int N, int V, int E>
static void launch_floor_col(float2* plane, int batch, int nx, hipStream_t s) {
  hipLaunchKernelGGL((k_floor_col<N, V, E>), dim3(nx / ColCfg<N, V, E>::COLS, batch),
                     dim3(ColCfg<N, V, E>::THREADS), 0, s, plane, nx);
}

typedef void (*floor_row_fn)(float2*, const float*, int, int, int, hipStream_t);
typedef void (*floor_col_fn)(float2*, int, int, hipStream_t);

template <int N>
static floor_row_fn pick_floor_row_n(int v) {
  constexpr int E0 = row_e_of(N, 0), E1 = row_e_of(N, 1);
  switch (v) {
    case 0: return launch_floor_row<N, 0, E0>;
    case 2: return launch_floor_row<N, 2, E0>;
    case 3: return launch_floor_row<N, 0, E1>;
    case 4: return launch_floor_row<N, 1, E1>;
    case 5: return launch_floor_row<N, 2, E1>;
    default: return launch_floor_row<N, 1, E0>;
  }
}

template <int N>
static floor_col_fn pick_floor_col_n(int v) {
  constexpr int E0 = col_e_of(N, 0), E1 = col_e_of(N, 1);
  switch (v) {
    case 0: return launch_floor_col<N, 0, E0>;
    case 2: return launch_floor_col<N, 2, E0>;
    case 3: return launch_floor_col<N, 3, E0>;
    case 4: return launch_floor_col<N, 4, E0>;
    case 5: return launch_floor_col<N, 0, E1>;
    case 6: return launch_floor_col<N, 1, E1>;
    case 7: return launch_floor_col<N, 2, E1>;
    case 8: return launch_floor_col<N, 3, E1>;
    case 9: return launch_floor_col<N, 4, E1>;
    default: return launch_floor_col<N, 1, E0>;
  }
}

extern "C" int bpmh_floor_steps(bpmh_plan plan, int batch, int nz, float* plane_, const float* dn,
                                long long dn_numel, void* stream) {
  if (!plan || !plane_ || !dn) return bpmh_set_error(BPMH_ERR_INVALID, "null argument");
  if (batch < 1 || nz < 1 || batch > 65535 || (long long)batch * plan->ny > (1LL << 30))
    return bpmh_set_error(BPMH_ERR_INVALID, "bad batch or depth");
  const int nx = plan->nx, ny = plan->ny;
  if (dn_numel < (long long)nz * nx * ny)
    return bpmh_set_error(BPMH_ERR_INVALID, "dn holds fewer than nz planes");
  float2* plane = (float2*)plane_;
  const int cv = col_variant_for(plan->col_v, plane);
  floor_row_fn fr = nullptr;
  floor_col_fn fc = nullptr;
#define FLOOR_CASE(N)                          \
  case N:                                      \
    fr = pick_floor_row_n<N>(plan->row_v);     \
    break;
  switch (nx) { BPMH_SIZES(FLOOR_CASE) }
#undef FLOOR_CASE
#define FLOOR_CASE(N)                  \
  case N:                              \
    fc = pick_floor_col_n<N>(cv);      \
    break;
  switch (ny) { BPMH_SIZES(FLOOR_CASE) }
#undef FLOOR_CASE
  if (!fr || !fc) return bpmh_set_error(BPMH_ERR_UNSUPPORTED, "grid size");
  hipStream_t s = (hipStream_t)stream;
  for (int z = 1; z < nz; ++z) {
    fc(plane, batch, nx, s);
    fr(plane, dn, batch, ny, z, s);
  }
  HIP_TRY(hipGetLastError());
  return BPMH_OK;
}

extern "C" int bpmh_propagate(bpmh_plan plan, int batch, int nz, float* plane, const void* dn,
                              int dn_kind, long long dn_batch_stride, long long dn_numel,
                              int mean_mode, const float* H, const float* n_eff, const float* dn0,
                              float n0, float k0, float dx, float dy, float dz, float unit_k,
                              int absorb, float* out_field, float* out_intens,
                              long long out_numel, float* out_sum, const float* scat_w,
                              const float* g_w, const float* plain, float* out_scat,
                              void* stream) {
  PropCall c;
  c.batch = batch;
  c.nz = nz;
  c.plane = (float2*)plane;
  c.dn = dn;
  c.dn_kind = dn_kind;
  c.dn_bstride = dn_batch_stride;
  c.dn_numel = dn_numel;
  c.mean_mode = mean_mode;
  c.H = (const float2*)H;
  c.n_eff = n_eff;
  c.dn0 = dn0;
  c.n0 = n0;
  c.k0 = k0;
  c.dx = dx;
  c.dy = dy;
  c.dz = dz;
  c.unit_k = unit_k;
  c.absorb = absorb;
  c.out_field = (float2*)out_field;
  c.out_intens = out_intens;
  c.out_numel = out_numel;
  c.out_sum = out_sum;
  c.scat_w = scat_w;
  c.g_w = g_w;
  c.plain = (const float2*)plain;
  c.out_scat = out_scat;
  int rc = prop_validate(plan, c);
  if (rc) return rc;
  rc = plan_upload(plan);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  StepScratch sc = {nullptr, nullptr, nullptr};
  const size_t sb = prop_scratch_bytes(plan, c);
  if (sb) {
    void* scr = nullptr;
    rc = plan_scratch(plan, sb, &scr);
    if (rc) return rc;
    sc = prop_scratch_carve(plan, c, scr);
  }
  return prop_run(plan, c, sc, s, true);
}

// ----------------------------------------------------------------------------
// ragged batch: groups of fields with their own dn volumes and depths
// ----------------------------------------------------------------------------
extern "C" int bpmh_propagate_groups(bpmh_plan plan, int n_groups, const int* items,
                                     const int* nzs, float* plane_, const void* dn, int dn_kind,
                                     long long dn_group_stride, long long dn_numel,
                                     const float* H_, float unit_k, int absorb, float* out_sum,
                                     void** group_done, void* stream) {
  float2* plane = (float2*)plane_;
  const float2* H = (const float2*)H_;
  char b[256];
  if (!plan) return bpmh_set_error(BPMH_ERR_INVALID, "null plan");
  if (!plane || !H) return bpmh_set_error(BPMH_ERR_INVALID, "null plane or propagator");
  if (n_groups < 1 || n_groups > BPMH_MAX_GROUPS || !items || !nzs)
    return bpmh_set_error(BPMH_ERR_INVALID, "1 to 16 groups with item and depth arrays");
  if (dn_kind != BPMH_DN_NONE && dn_kind != BPMH_DN_F32 && dn_kind != BPMH_DN_F16 &&
      dn_kind != BPMH_DN_C64)
    return bpmh_set_error(BPMH_ERR_INVALID, "unknown dn kind");
  if (dn_kind != BPMH_DN_NONE && !dn) return bpmh_set_error(BPMH_ERR_INVALID, "null dn");
  if (dn_group_stride < 0 || absorb < 0)
    return bpmh_set_error(BPMH_ERR_INVALID, "negative stride or absorbing width");
  const int nx = plan->nx, ny = plan->ny;
  const long long n = (long long)nx * ny;
  int total = 0;
  for (int g = 0; g < n_groups; ++g) {
    if (items[g] < 1 || nzs[g] < 1)
      return bpmh_set_error(BPMH_ERR_INVALID, "every group needs >= 1 item and >= 1 plane");
    if (g > 0 && nzs[g] > nzs[g - 1])
      return bpmh_set_error(BPMH_ERR_INVALID, "groups must come in order of non-increasing depth");
    if (dn_kind != BPMH_DN_NONE &&
        (long long)g * dn_group_stride + (long long)nzs[g] * n > dn_numel) {
      snprintf(b, sizeof b, "dn holds %lld elements; group %d reads %d planes at offset %lld",
               dn_numel, g, nzs[g], (long long)g * dn_group_stride);
      return bpmh_set_error(BPMH_ERR_INVALID, b);
    }
    total += items[g];
    if (total > 65535) return bpmh_set_error(BPMH_ERR_INVALID, "batch too large");
  }
  if ((long long)total * ny > (1LL << 30)) return bpmh_set_error(BPMH_ERR_INVALID, "batch too large");
  int rc = plan_upload(plan);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  const row_fn row = pick_row(nx, plan->row_v);
  const int cv = col_variant_for(plan->col_v, plane);
  const col_fn col = pick_col<0>(ny, cv);

  RowArgs ra;
  memset(&ra, 0, sizeof ra);
  ra.plane = plane;
  ra.tw = plan->tw_x[plan->row_v / kRowShapes];
  ra.dn = dn;
  ra.dn_bstride = dn_group_stride;
  ra.dn_kind = dn_kind;
  ra.ngroups = n_groups;
  for (int g = 0, e = 0; g < n_groups; ++g) ra.gend[g] = (e += items[g]);
  ra.ny = ny;
  ra.nz = nzs[0];
  ra.unit_k = unit_k;
  ra.inv_n = 1.0f / ((float)nx * (float)ny);
  ra.absorb = absorb;
  ra.acc = out_sum;

  ColArgs ca;
  memset(&ca, 0, sizeof ca);
  ca.plane = plane;
  ca.tw = plan->tw_y[cv / kColShapes];
  ca.nx = nx;
  ca.nz = nzs[0];
  ca.hmode = 0;
  ca.H = H;

  // items still propagating at plane z (and, for the forward transform, past
  // it): a prefix, the groups being in order of non-increasing depth
  auto items_deeper = [&](int z) {
    int c = 0;
    for (int g = 0; g < n_groups; ++g)
      if (nzs[g] > z) c += items[g];
    return c;
  };
  auto mark_done = [&](int z) -> int {  // groups whose last plane is z
    if (!group_done) return BPMH_OK;
    for (int g = 0; g < n_groups; ++g)
      if (nzs[g] - 1 == z && group_done[g]) HIP_TRY(hipEventRecord((hipEvent_t)group_done[g], s));
    return BPMH_OK;
  };
  ra.z = 0;
  ra.acc_init = 1;
  ra.inverse = 0;
  ra.screen = 0;
  ra.fwd_items = items_deeper(1);
  row(ra, total, s);
  rc = mark_done(0);
  if (rc) return rc;
  ra.acc_init = 0;
  ra.inverse = 1;
  ra.screen = dn_kind != BPMH_DN_NONE;
  for (int z = 1; z < nzs[0]; ++z) {
    const int active = items_deeper(z);
    ca.z = z;
    col(ca, active, s);
    ra.z = z;
    ra.fwd_items = items_deeper(z + 1);
    row(ra, active, s);
    rc = mark_done(z);
    if (rc) return rc;
  }
  HIP_TRY(hipGetLastError());
  return BPMH_OK;
}

// ----------------------------------------------------------------------------
// allocation-free variant of bpmh_propagate
// ----------------------------------------------------------------------------
// Its step scratch (weighted-mean and scattering partials, weighted dn0) is
// this fixed device array instead of plan memory.  Each GPU holds its own
// copy for the life of the process.  In host code `scratch` names a
// host-side stand-in, not device memory: the device address comes from
// hipGetSymbolAddress, once per device, and only that address is used.
constexpr size_t kNoAllocScratchBytes = (size_t)16 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static std::mutex g_noalloc_mu;
static void* g_noalloc_ptr[64] = {};  // per device

// Device address of this device's `scratch`, resolved on the first call on
// the device (with that device current) and kept.
static int noalloc_scratch(void** out) {
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return bpmh_set_error(BPMH_ERR_UNSUPPORTED, "device index above 63");
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      char b[256];
      snprintf(b, sizeof b, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
               hipGetErrorString(e));
      return bpmh_set_error(BPMH_ERR_HIP, b);
    }
    g_noalloc_ptr[dev] = p;
  }
  *out = g_noalloc_ptr[dev];
  return BPMH_OK;
}

extern "C" int bpmh_propagate_noalloc(bpmh_plan plan, int batch, int nz, float* plane,
                                      const void* dn, int dn_kind, long long dn_batch_stride,
                                      long long dn_numel, int mean_mode, const float* H,
                                      const float* n_eff, const float* dn0, float n0, float k0,
                                      float dx, float dy, float dz, float unit_k, int absorb,
                                      float* out_field, float* out_intens, long long out_numel,
                                      float* out_sum, const float* scat_w, const float* g_w,
                                      const float* plain, float* out_scat, void* stream) {
  PropCall c;
  c.batch = batch;
  c.nz = nz;
  c.plane = (float2*)plane;
  c.dn = dn;
  c.dn_kind = dn_kind;
  c.dn_bstride = dn_batch_stride;
  c.dn_numel = dn_numel;
  c.mean_mode = mean_mode;
  c.H = (const float2*)H;
  c.n_eff = n_eff;
  c.dn0 = dn0;
  c.n0 = n0;
  c.k0 = k0;
  c.dx = dx;
  c.dy = dy;
  c.dz = dz;
  c.unit_k = unit_k;
  c.absorb = absorb;
  c.out_field = (float2*)out_field;
  c.out_intens = out_intens;
  c.out_numel = out_numel;
  c.out_sum = out_sum;
  c.scat_w = scat_w;
  c.g_w = g_w;
  c.plain = (const float2*)plain;
  c.out_scat = out_scat;
  int rc = prop_validate(plan, c);
  if (rc) return rc;
  if (!plan->tw_x[0])
    return bpmh_set_error(BPMH_ERR_INVALID,
                          "plan has no device tables: create it with a device present");
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  if (dev != plan->device) return bpmh_set_error(BPMH_ERR_INVALID, "plan belongs to another device");
  StepScratch sc = {nullptr, nullptr, nullptr};
  const size_t sb = prop_scratch_bytes(plan, c);
  if (sb > kNoAllocScratchBytes) {
    char b[256];
    snprintf(b, sizeof b, "this call needs %zu bytes of scratch; the fixed scratch holds %zu "
             "(fewer items or planes with scattering sums)", sb, kNoAllocScratchBytes);
    return bpmh_set_error(BPMH_ERR_INVALID, b);
  }
  if (sb) {
    void* p = nullptr;
    rc = noalloc_scratch(&p);
    if (rc) return rc;
    sc = prop_scratch_carve(plan, c, p);
  }
  return prop_run(plan, c, sc, (hipStream_t)stream, false);
}

// ----------------------------------------------------------------------------
// adjoint (gradient of a loss on the last plane with respect to dn)
// ----------------------------------------------------------------------------
// grad[i] += sum_b gitem[b][i] in item order (deterministic); n counts floats
// (one per voxel for real dn, two for complex dn)
__global__ void k_grad_reduce(const float* __restrict__ gitem, float* __restrict__ grad,
                              int batch, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = grad[i];
  for (int b = 0; b < batch; ++b) s += gitem[(size_t)b * n + i];
  grad[i] = s;
}

extern "C" int bpmh_adjoint_segment(bpmh_plan plan, int batch, int z_lo, int z_hi, float* g_,
                                    const float* u_seg_, long long u_seg_numel, const void* dn,
                                    int dn_kind, long long dn_numel, const float* H_,
                                    float unit_k, int absorb, float* grad_dn, float* gitem,
                                    void* stream) {
  float2* g = (float2*)g_;
  const float2* u_seg = (const float2*)u_seg_;
  const float2* H = (const float2*)H_;
  if (!plan || !g || !u_seg || !dn || !H || !grad_dn || !gitem)
    return bpmh_set_error(BPMH_ERR_INVALID, "null argument");
  if (dn_kind != BPMH_DN_F32 && dn_kind != BPMH_DN_C64)
    return bpmh_set_error(BPMH_ERR_UNSUPPORTED, "the adjoint takes float32 or complex64 dn");
  if (batch < 1 || z_lo < 0 || z_hi < z_lo) return bpmh_set_error(BPMH_ERR_INVALID, "bad range");
  if (batch > 65535 || (long long)batch * plan->ny > (1LL << 30))
    return bpmh_set_error(BPMH_ERR_INVALID, "batch too large");
  if (absorb < 0) return bpmh_set_error(BPMH_ERR_INVALID, "absorbing width must be >= 0");
  const int nx = plan->nx, ny = plan->ny;
  const int nseg = z_hi - z_lo + 1;
  const long long n2 = (long long)nx * ny;
  char b[256];
  if (dn_numel < (long long)(z_hi + 1) * n2) {
    snprintf(b, sizeof b, "dn and grad_dn hold %lld elements; planes up to %d need %lld",
             dn_numel, z_hi, (long long)(z_hi + 1) * n2);
    return bpmh_set_error(BPMH_ERR_INVALID, b);
  }
  if (u_seg_numel < (long long)batch * nseg * n2) {
    snprintf(b, sizeof b, "u_seg holds %lld elements; %d item(s) x %d planes need %lld",
             u_seg_numel, batch, nseg, (long long)batch * nseg * n2);
    return bpmh_set_error(BPMH_ERR_INVALID, b);
  }
  if (z_hi == z_lo) return BPMH_OK;
  const int urc = plan_upload(plan);
  if (urc) return urc;
  hipStream_t s = (hipStream_t)stream;
  const row_fn row = pick_row(nx, plan->row_v);
  const int cv = col_variant_for(plan->col_v, g);
  const col_fn col0 = pick_col<0>(ny, cv);
  const int comps = dn_kind == BPMH_DN_C64 ? 2 : 1;  // floats of gradient per voxel

  RowArgs ra;
  memset(&ra, 0, sizeof ra);
  ra.plane = g;
  ra.tw = plan->tw_x[plan->row_v / kRowShapes];
  ra.dn = dn;
  ra.dn_kind = dn_kind;
  ra.ny = ny;
  ra.nz = z_hi + 1;
  ra.unit_k = unit_k;
  ra.inv_n = 1.0f / ((float)nx * (float)ny);
  ra.absorb = absorb;
  ra.adjoint = 1;
  ra.useg = u_seg;
  ra.useg_bstride = (long long)nseg * ny * nx;
  ra.gitem = gitem;

  ColArgs ca;
  memset(&ca, 0, sizeof ca);
  ca.plane = g;
  ca.tw = plan->tw_y[cv / kColShapes];
  ca.nx = nx;
  ca.hmode = 0;
  ca.H = H;
  ca.conj_h = 1;

  const size_t n = (size_t)n2 * comps;
  for (int z = z_hi; z > z_lo; --z) {
    ra.z = z;
    ra.useg_z = z - z_lo;
    ra.inverse = z != z_hi;
    ra.screen = 1;
    ra.fwd_items = batch;
    row(ra, batch, s);
    hipLaunchKernelGGL(k_grad_reduce, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, gitem,
                       grad_dn + (size_t)z * n, batch, n);
    col0(ca, batch, s);
  }
  // back to real space: the gradient with respect to u[z_lo]
  ra.adjoint = 0;
  ra.screen = 0;
  ra.inverse = 1;
  ra.fwd_items = 0;
  row(ra, batch, s);
  HIP_TRY(hipGetLastError());
  return BPMH_OK;
}

// ----------------------------------------------------------------------------
// mean-corrected (weighted or simple) segment of the differentiable form
// ----------------------------------------------------------------------------
// out[r] = sum of row r (nx floats) of g; one wavefront per row, fixed order
__global__ void __launch_bounds__(256) k_row_sums(const float* __restrict__ g,
                                                  float* __restrict__ out, int rows, int nx) {
  const int r = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (r >= rows) return;  // whole wavefront
  const float* p = g + (size_t)r * nx;
  float sum = 0.f;
  for (int i = lane; i < nx; i += 64) sum += p[i];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o, 64);
  if (lane == 0) out[r] = sum;
}

__global__ void k_fill_zero(float* __restrict__ p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

// Workspace of bpmh_mean_segment, in floats (each part 256-byte aligned).
struct WsegLayout {
  size_t wpart, dn0, fplane, useg, kseg, recc, recw, s1, s2, gitem, gd, cg, total;
};

static WsegLayout wseg_layout(int nx, int ny, int batch, int nseg, bool adjoint) {
  WsegLayout L;
  memset(&L, 0, sizeof L);
  size_t o = 0;
  auto take = [&o](size_t n) {
    const size_t r = o;
    o += (n + 63) & ~(size_t)63;
    return r;
  };
  const size_t n2 = (size_t)nx * ny, B = (size_t)batch, S = (size_t)nseg;
  L.wpart = take(2 * B * nx);
  L.dn0 = take(B);
  if (adjoint) {
    L.fplane = take(2 * B * n2);
    L.useg = take(2 * B * S * n2);
    L.kseg = take(2 * B * (S - 1) * n2);
    L.recc = take((S - 1) * B);
    L.recw = take((S - 1) * B);
    L.s1 = take(B * ny);
    L.s2 = take(B * nx);
    L.gitem = take(B * n2);
    L.gd = take(B * n2);
    L.cg = take(B);
  }
  L.total = o;
  return L;
}

// grad[i] += (sum_b cg[b]) / n over one plane: dL/ddn0 of a simple-mean step
// spread over the plane it is the mean of (items summed in order)
__global__ void k_add_mean_grad(float* __restrict__ grad, const float* __restrict__ cg, int batch,
                                size_t n, float inv_n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float c = 0.f;
  for (int b = 0; b < batch; ++b) c += cg[b];
  grad[i] += c * inv_n;
}

extern "C" int bpmh_mean_segment_work(bpmh_plan plan, int batch, int z_lo, int z_hi,
                                      int adjoint, long long* numel) {
  if (!plan || !numel) return bpmh_set_error(BPMH_ERR_INVALID, "null argument");
  if (batch < 1 || z_lo < 0 || z_hi < z_lo) return bpmh_set_error(BPMH_ERR_INVALID, "bad range");
  *numel = (long long)wseg_layout(plan->nx, plan->ny, batch, z_hi - z_lo + 1, adjoint != 0).total;
  return BPMH_OK;
}

extern "C" int bpmh_mean_segment(bpmh_plan plan, int mean_mode, int batch, int z_lo, int z_hi,
                                 float* plane_, const float* dn, long long dn_numel,
                                 const float* n_eff, const float* dn0_arr, float n0, float k0,
                                 float dx, float dy, float dz, float unit_k, int absorb, float* g_,
                                 float* grad_dn, float* work, long long work_numel, void* stream) {
  if (!plan || !plane_ || !dn || !work) return bpmh_set_error(BPMH_ERR_INVALID, "null argument");
  if (mean_mode != BPMH_MEAN_SIMPLE && mean_mode != BPMH_MEAN_WEIGHTED)
    return bpmh_set_error(BPMH_ERR_INVALID, "mean_mode must be simple or weighted");
  const bool simple = mean_mode == BPMH_MEAN_SIMPLE;
  if (simple && (!n_eff || !dn0_arr))
    return bpmh_set_error(BPMH_ERR_INVALID, "mean_mode simple needs n_eff and dn0 arrays");
  const bool adjoint = g_ != nullptr;
  if (adjoint && !grad_dn) return bpmh_set_error(BPMH_ERR_INVALID, "the adjoint needs grad_dn");
  if (batch < 1 || z_lo < 0 || z_hi < z_lo) return bpmh_set_error(BPMH_ERR_INVALID, "bad range");
  if (batch > 65535 || (long long)batch * plan->ny > (1LL << 30))
    return bpmh_set_error(BPMH_ERR_INVALID, "batch too large");
  if (absorb < 0) return bpmh_set_error(BPMH_ERR_INVALID, "absorbing width must be >= 0");
  const int nx = plan->nx, ny = plan->ny;
  const int nseg = z_hi - z_lo + 1;
  const long long n2 = (long long)nx * ny;
  char msg[256];
  if (dn_numel < (long long)(z_hi + 1) * n2) {
    snprintf(msg, sizeof msg, "dn holds %lld elements; planes up to %d need %lld", dn_numel, z_hi,
             (long long)(z_hi + 1) * n2);
    return bpmh_set_error(BPMH_ERR_INVALID, msg);
  }
  const WsegLayout L = wseg_layout(nx, ny, batch, nseg, adjoint);
  if (work_numel < (long long)L.total) {
    snprintf(msg, sizeof msg, "work holds %lld floats; this segment needs %lld", work_numel,
             (long long)L.total);
    return bpmh_set_error(BPMH_ERR_INVALID, msg);
  }
  if (nseg == 1) return BPMH_OK;
  const int urc = plan_upload(plan);
  if (urc) return urc;
  hipStream_t s = (hipStream_t)stream;
  const row_fn row = pick_row(nx, plan->row_v);
  float2* plane = (float2*)plane_;
  float2* fp = adjoint ? (float2*)(work + L.fplane) : plane;
  const int cv = col_variant_for(plan->col_v, fp);
  const col_fn col0 = simple ? pick_col<5>(ny, cv) : pick_col<0>(ny, cv);
  const col_fn colA = pick_col<1>(ny, cv);
  const col_fn colB = pick_col<2>(ny, cv);
  const size_t B = (size_t)batch, S = (size_t)nseg;
  float2* useg = (float2*)(work + L.useg);
  float2* kseg = (float2*)(work + L.kseg);
  if (adjoint)
    HIP_TRY(hipMemcpyAsync(fp, plane, B * n2 * sizeof(float2), hipMemcpyDeviceToDevice, s));

  // forward z_lo -> z_hi (recorded when the adjoint follows); the weights
  // read slice 0 of the whole volume, the screens dn[z_lo + z]
  RowArgs ra;
  memset(&ra, 0, sizeof ra);
  ra.plane = fp;
  ra.tw = plan->tw_x[plan->row_v / kRowShapes];
  ra.dn = dn + (size_t)z_lo * n2;
  ra.dn_kind = BPMH_DN_F32;
  ra.ny = ny;
  ra.nz = nseg;
  ra.unit_k = unit_k;
  ra.inv_n = 1.0f / ((float)nx * (float)ny);
  ra.absorb = absorb;
  ra.dn0_mode = mean_mode;
  ra.dn0_item = work + L.dn0;
  ra.dn0_arr = simple ? dn0_arr + z_lo : nullptr;  // per plane, indexed by the segment's z
  ra.out_field = adjoint ? useg : nullptr;

  ColArgs ca;
  memset(&ca, 0, sizeof ca);
  ca.plane = fp;
  ca.tw = plan->tw_y[cv / kColShapes];
  ca.nx = nx;
  ca.nz = nseg;
  ca.hmode = mean_mode;
  ca.n_eff = simple ? n_eff + z_lo : nullptr;
  ca.n0 = n0;
  ca.k0 = k0;
  ca.dx = dx;
  ca.dy = dy;
  ca.dz = dz;
  ca.dn = dn;
  ca.dn_kind = BPMH_DN_F32;
  ca.wpart = (float2*)(work + L.wpart);
  ca.dn0_item = work + L.dn0;

  ra.z = 0;
  ra.fwd_items = batch;
  row(ra, batch, s);
  ra.inverse = 1;
  ra.screen = 1;
  for (int z = 1; z < nseg; ++z) {
    ca.z = z;
    if (adjoint) {
      ca.kout = kseg + (size_t)(z - 1) * n2;
      ca.kout_bstride = (long long)(S - 1) * n2;
      ca.rec_c = work + L.recc + (size_t)(z - 1) * B;
      ca.rec_w = work + L.recw + (size_t)(z - 1) * B;
    }
    if (simple) {
      col0(ca, batch, s);
    } else {
      colA(ca, batch, s);
      colB(ca, batch, s);
    }
    ra.z = z;
    ra.fwd_items = z < nseg - 1 ? batch : 0;
    row(ra, batch, s);
  }
  if (!adjoint) {
    HIP_TRY(hipGetLastError());
    return BPMH_OK;
  }

  // adjoint z_hi -> z_lo
  float2* g = (float2*)g_;
  float* gitem = work + L.gitem;
  float* gd = work + L.gd;
  const size_t nn = (size_t)n2;
  const unsigned fill_blocks = (unsigned)((B * nn + 255) / 256);
  hipLaunchKernelGGL(k_fill_zero, dim3(fill_blocks), dim3(256), 0, s, gd, B * nn);
  const col_fn col3 = pick_col_adj<3>(ny, cv);
  const col_fn col4 = pick_col_adj<4>(ny, cv);
  RowArgs rb = ra;
  rb.plane = g;
  rb.out_field = nullptr;
  rb.adjoint = 1;
  rb.useg = useg;
  rb.useg_bstride = (long long)S * n2;
  rb.gitem = gitem;
  rb.screen = 1;
  rb.fwd_items = batch;
  ColArgs cb;
  memset(&cb, 0, sizeof cb);
  cb.plane = g;
  cb.tw = ca.tw;
  cb.nx = nx;
  cb.nz = nseg;
  cb.hmode = mean_mode;
  cb.simple = simple;
  cb.n_eff = simple ? n_eff + z_lo : nullptr;
  cb.cg_out = work + L.cg;
  cb.conj_h = 1;
  cb.n0 = n0;
  cb.k0 = k0;
  cb.dx = dx;
  cb.dy = dy;
  cb.dz = dz;
  cb.dn = dn;
  cb.dn_kind = BPMH_DN_F32;
  cb.uk_bstride = (long long)(S - 1) * n2;
  cb.s2part = work + L.s2;
  cb.s1part = work + L.s1;
  cb.gd_item = gd;
  const unsigned red_blocks = (unsigned)((nn + 255) / 256);
  for (int z = nseg - 1; z >= 1; --z) {
    rb.z = z;
    rb.useg_z = z;
    rb.inverse = z != nseg - 1;
    if (!simple) rb.dn0_item = work + L.recc + (size_t)(z - 1) * B;
    row(rb, batch, s);
    hipLaunchKernelGGL(k_grad_reduce, dim3(red_blocks), dim3(256), 0, s, gitem,
                       grad_dn + (size_t)(z_lo + z) * nn, batch, nn);
    hipLaunchKernelGGL(k_row_sums, dim3((unsigned)((B * ny + 3) / 4)), dim3(256), 0, s, gitem,
                       work + L.s1, (int)(B * ny), nx);
    cb.z = z;
    cb.uk = kseg + (size_t)(z - 1) * n2;
    cb.cz = work + L.recc + (size_t)(z - 1) * B;
    cb.wz = work + L.recw + (size_t)(z - 1) * B;
    col3(cb, batch, s);
    col4(cb, batch, s);
    if (simple)  // dn0 of this step is the mean of plane z: dL/ddn0 / N on every voxel
      hipLaunchKernelGGL(k_add_mean_grad, dim3(red_blocks), dim3(256), 0, s,
                         grad_dn + (size_t)(z_lo + z) * nn, (const float*)(work + L.cg), batch, nn,
                         1.f / (float)nn);
  }
  rb.adjoint = 0;
  rb.screen = 0;
  rb.inverse = 1;
  rb.fwd_items = 0;
  row(rb, batch, s);
  // the weight slice is slice 0 of the volume
  if (!simple)
    hipLaunchKernelGGL(k_grad_reduce, dim3(red_blocks), dim3(256), 0, s, gd, grad_dn, batch, nn);
  HIP_TRY(hipGetLastError());
  return BPMH_OK;
}
