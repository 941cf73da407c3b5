// Refractive-index volumes from spheres, generated on the device.
//
// k_fill_dn_coated: the coated-sphere volume of bpm/bpm/kernels/bpm_3d_spheres.cl:31-74
//   (voxel r = (i dx, j dy, k dz); dmin = distance to the nearest centre
//   among the spheres whose z lies within rad2 of r.z (centres sorted by z);
//   dn = dmin < rad1 ? dn1 : dmin < rad2 ? dn2 : 0).
// k_spheres_sum: sum over spheres of dn_s [|r - c_s| < r_s], the volumes of
//   the benchmark workloads; distances are tested in float64 and the float32
//   sum follows the sphere order, so the result equals the host generator
//   bpm_amd.workloads.spheres_dn bit for bit.
// Both are HBM-write-bound (4 B per voxel written).
#include "common.hpp"
#include "bpm_hip.h"

int bpmh_set_error(int code, const char* msg);

__global__ void k_fill_dn_coated(float* dn, int nx, int ny, int nz, const float* pts, int np,
                                 float dx, float dy, float dz, float dn1, float rad1, float dn2,
                                 float rad2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  const int k = blockIdx.z;
  if (i >= nx) return;
  const float rx = f_mul(dx, (float)i), ry = f_mul(dy, (float)j), rz = f_mul(dz, (float)k);
  // first centre with z >= rz - rad2 and last with z <= rz + rad2 (the
  // reference's two linear scans, as binary searches; clamped the same way)
  const float zlo = f_sub(rz, rad2), zhi = f_add(rz, rad2);
  int lo = 0, hi = np - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (pts[3 * mid + 2] < zlo) lo = mid + 1; else hi = mid;
  }
  const int ind1 = lo;
  lo = 0;
  hi = np - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (pts[3 * mid + 2] > zhi) hi = mid - 1; else lo = mid;
  }
  const int ind2 = lo;
  float dmin = f_mul(1000000000.f, rad2);
  for (int n = ind1; n <= ind2; ++n) {
    const float ex = f_sub(rx, pts[3 * n]);
    const float ey = f_sub(ry, pts[3 * n + 1]);
    const float ez = f_sub(rz, pts[3 * n + 2]);
    const float d = __fsqrt_rn(f_add(f_add(f_mul(ex, ex), f_mul(ey, ey)), f_mul(ez, ez)));
    dmin = d < dmin ? d : dmin;
  }
  dn[((size_t)k * ny + j) * nx + i] = dmin < rad2 ? (dmin < rad1 ? dn1 : dn2) : 0.f;
}

// spheres: n x 5 float64 (cx, cy, cz, r, dn) in micrometres
__global__ void k_spheres_sum(float* dn, int nx, int ny, int nz, const double* sph, int n,
                              double dx, double dy, double dz) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  const int k = blockIdx.z;
  if (i >= nx) return;
  const double x = d_mul((double)i, dx), y = d_mul((double)j, dy), z = d_mul((double)k, dz);
  float acc = 0.f;
  for (int s = 0; s < n; ++s) {
    const double* p = sph + 5 * s;
    const double ez = d_sub(z, p[2]);
    const double r = p[3];
    if (fabs(ez) >= r) continue;
    const double ey = d_sub(y, p[1]);
    if (fabs(ey) >= r) continue;
    const double ex = d_sub(x, p[0]);
    const double d2 = d_add(d_add(d_mul(ez, ez), d_mul(ey, ey)), d_mul(ex, ex));
    if (d2 < d_mul(r, r)) acc += (float)p[4];
  }
  dn[((size_t)k * ny + j) * nx + i] = acc;
}

extern "C" int bpmh_fill_dn_coated(int nx, int ny, int nz, float dx, float dy, float dz,
                                   const float* points, int n_points, float dn1, float rad1,
                                   float dn2, float rad2, float* dn, void* stream) {
  if (nx < 1 || ny < 1 || nz < 1 || ny > 65535 || nz > 65535)
    return bpmh_set_error(BPMH_ERR_INVALID, "bad grid");
  if (n_points < 1 || !points || !dn) return bpmh_set_error(BPMH_ERR_INVALID, "need points and dn");
  dim3 grid((nx + 127) / 128, ny, nz);
  hipLaunchKernelGGL(k_fill_dn_coated, grid, dim3(128), 0, (hipStream_t)stream, dn, nx, ny, nz,
                     points, n_points, dx, dy, dz, dn1, rad1, dn2, rad2);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return bpmh_set_error(BPMH_ERR_HIP, hipGetErrorString(e));
  return BPMH_OK;
}

extern "C" int bpmh_spheres_dn(int nx, int ny, int nz, double dx, double dy, double dz,
                               const double* spheres, int n_spheres, float* dn, void* stream) {
  if (nx < 1 || ny < 1 || nz < 1 || ny > 65535 || nz > 65535)
    return bpmh_set_error(BPMH_ERR_INVALID, "bad grid");
  if (n_spheres < 0 || (n_spheres > 0 && !spheres) || !dn)
    return bpmh_set_error(BPMH_ERR_INVALID, "bad sphere table or output");
  dim3 grid((nx + 127) / 128, ny, nz);
  hipLaunchKernelGGL(k_spheres_sum, grid, dim3(128), 0, (hipStream_t)stream, dn, nx, ny, nz,
                     spheres, n_spheres, dx, dy, dz);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return bpmh_set_error(BPMH_ERR_HIP, hipGetErrorString(e));
  return BPMH_OK;
}
