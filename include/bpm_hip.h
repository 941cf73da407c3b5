/* C ABI of libbpmhip.so: split-step BPM and vectorial Debye-Wolf on gfx950.
 *
 * Every entry point returns BPMH_OK (0) or a negative BPMH_ERR_* code and
 * never throws; bpmh_last_error() holds a thread-local message for the last
 * failure.  All array arguments are DEVICE pointers owned by the caller
 * (C-order, x fastest; a complex64 array is a float* of interleaved
 * (re, im) pairs); work is enqueued on
 * the caller's hipStream_t (passed as void*, NULL = default stream) and is
 * asynchronous.  No entry point takes a host array except where stated.
 * The propagation entry points take the element counts of their
 * variable-length device arrays (the dn volume, the output volumes, the
 * recomputed forward planes) and return BPMH_ERR_INVALID, before launching
 * anything, when an array is shorter than the call would read or write.
 * Arrays whose length follows from the plan and the sizes (planes, H,
 * per-plane arrays) must hold what the declaration says.
 *
 * The reference binds its device work through gputools/pyopencl
 * (OCLProgram.run_kernel, fft, OCLReductionKernel); each declaration below
 * names the reference interface it replaces.
 */
#ifndef BPM_HIP_H
#define BPM_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define BPMH_VERSION 201

#define BPMH_OK 0
#define BPMH_ERR_INVALID -1      /* bad argument  (Python: ValueError) */
#define BPMH_ERR_UNSUPPORTED -2  /* unsupported   (Python: NotImplementedError) */
#define BPMH_ERR_HIP -3          /* HIP runtime error */
#define BPMH_ERR_NOMEM -4        /* device allocation failed */

/* dn kinds (bpm_3d.py:198-214) */
#define BPMH_DN_NONE 0
#define BPMH_DN_F32 1
#define BPMH_DN_F16 2
#define BPMH_DN_C64 3

/* mean modes (bpm_3d.py:54,120,286-317) */
#define BPMH_MEAN_IGNORE 0   /* static propagator table H */
#define BPMH_MEAN_SIMPLE 1   /* H(n_eff[z]) and screen dn - dn0[z], per-plane arrays */
#define BPMH_MEAN_WEIGHTED 2 /* dn0 = sum|U| dn_slice0 / sum|U| on device each step */

/* propagator table kinds */
#define BPMH_H_F32 0  /* compute_propagator, bpm/bpm/kernels/bpm_3d_kernels.cl:408-432 */
#define BPMH_H_F64 1  /* host float64 H, bpm/bpmclass/bpm3d.py:88-107 (rounded to complex64) */

typedef struct bpmh_plan_s* bpmh_plan;

/* Library version (BPMH_VERSION). */
int bpmh_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char* bpmh_last_error(void);

/* Plan for an (ny, nx) grid: FFT twiddles and step scratch.
 * Replaces gputools fft_plan((Ny,Nx)) (bpm/bpm/bpm_3d.py:188,
 * bpm/bpmclass/_bpm3d_ocl.py:43).  nx, ny: powers of two in [16, 4096]. */
int bpmh_plan_create(int nx, int ny, bpmh_plan* out);
int bpmh_plan_destroy(bpmh_plan plan);

/* Measurement hooks (the reference only prints wall-clock StopWatch times,
 * bpm/utils/utils.py:7-37, bpm/bpm/bpm_3d.py:140-142,384-386).  With timing
 * enabled, bpmh_propagate records a HIP event pair around every kernel it
 * launches; bpmh_plan_kernel_times waits for them and returns, per kernel
 * kind (0 row pass, 1 column pass, 2/3 weighted column passes A/B), the
 * summed milliseconds and launch counts since the previous read-out. */
int bpmh_plan_set_timing(bpmh_plan plan, int enable);
int bpmh_plan_kernel_times(bpmh_plan plan, float* ms, int* counts);

/* Graph mode: bpmh_propagate captures its whole launch sequence (2 or 3
 * kernels per z-step) into a HIP graph the first time it sees an argument
 * list and replays it (one hipGraphLaunch) whenever that argument list
 * recurs.  A plan keeps the 4 most recently used executables; one is
 * destroyed (on eviction, on bpmh_plan_set_graph(plan, 0) and at
 * bpmh_plan_destroy) only after its latest replay has completed.  Replaces
 * the per-kernel Python->pyopencl enqueue of the reference loop
 * (bpm/bpm/bpm_3d.py:283-382).  Ignored while timing is on. */
int bpmh_plan_set_graph(bpmh_plan plan, int enable);

/* Launch-shape and radix-schedule variants of the row and column passes,
 * chosen per grid size and batch by the autotuner (bpm_amd/tuning.py).
 * Row variant = V + 3*s, column variant = V + 5*s, where s selects the
 * radix schedule: 0 the size's default elements per thread (rows 4 up to
 * 1024 points, 8 above; columns 8, 4 up to 32 points), 1 the other one (8
 * instead of 4 from 64 points on, 4 instead of 8), e.g. a 1024-point line as
 * 4.4.4.4.4 or 2.8.8.8.  V: 0, 1, 2 half, the default, twice the lines /
 * columns per work-group; column V = 3 the work-group of 0 with two adjacent
 * columns per thread (tiles twice as wide, 16-byte accesses; 8-byte aligned
 * planes run V = 0), V = 4 the same with each column's loads and stores
 * issued separately (the first column transformed while the second one's
 * loads are in flight); V = 3, 4 apply to columns of 256-1024 points and are
 * V = 0 elsewhere.  Variants of one radix schedule give the same bits; the
 * two schedules round differently (both within the parity bar).
 * get_variant also reports the lines per row work-group, columns per column
 * tile and elements per thread of each pass.  No reference counterpart
 * (gputools picks its own FFT plans, bpm/bpm/bpm_3d.py:188). */
/* Memory floor of nz-1 z-steps on this plan (measurement only): per step
 * the column pass's and then the row pass's global loads and stores in the
 * plan's current launch shapes -- the plane read and written by both, one
 * float32 dn slice (nz, ny, nx; dn_numel elements) read by the row pattern,
 * 36 B per voxel-step -- with no transforms; the plane's values are
 * scrambled.  A propagation's step time over this one's is the cost of the
 * transforms on top of the step's data movement.  No reference
 * counterpart. */
int bpmh_floor_steps(bpmh_plan plan, int batch, int nz, float* plane, const float* dn,
                     long long dn_numel, void* stream);

int bpmh_plan_set_variant(bpmh_plan plan, int row_variant, int col_variant);
int bpmh_plan_get_variant(bpmh_plan plan, int* row_variant, int* col_variant,
                          int* lines_per_group, int* cols_per_group, int* row_elems,
                          int* col_elems);

/* Fill H (ny, nx) complex64 with the angular-spectrum propagator.
 * kind BPMH_H_F32 replaces the compute_propagator kernel launch
 * (bpm/bpm/bpm_3d.py:192-196, bpm_3d_kernels.cl:408-432), all arguments used
 * as float32; kind BPMH_H_F64 replaces the host numpy H of the class path
 * (bpm/bpmclass/bpm3d.py:88-107) uploaded at bpm/bpmclass/_bpm3d_ocl.py:46. */
int bpmh_propagator_table(int nx, int ny, int kind, double dx, double dy, double dz,
                          double n0, double lam, float* H, void* stream);

/* Run the whole split-step loop over nz planes (nz-1 steps) for `batch`
 * fields that share the grid (e.g. illumination angles).  Replaces the loop
 * body of bpm/bpm/bpm_3d.py:283-382 (fft, mult / mult_propagator /
 * mult_propagator_buff, dn_sum/u_sum reductions, scattering reductions, ifft,
 * mult_dn* kernels, copy_intens) and of bpm/bpmclass/_bpm3d_ocl.py:100-126.
 *
 *  plane      (batch, ny, nx) c64: u0 on entry, last plane u[nz-1] on exit
 *  dn         (nz, ny, nx) of dn_kind per item, item b at dn + b*dn_batch_stride
 *             elements (0: all items share one volume); NULL for BPMH_DN_NONE
 *  dn_numel   elements dn holds: at least (batch-1)*dn_batch_stride + nz*ny*nx
 *  mean_mode  BPMH_MEAN_*; IGNORE uses H (a table from bpmh_propagator_table;
 *             only its quadrant j <= ny/2, i <= nx/2 is read, the rest
 *             follows from H[j][i] = H[(ny-j)%ny][(nx-i)%nx]),
 *             SIMPLE uses n_eff[z] (float32 n0+mean(dn[z])) and dn0[z],
 *             WEIGHTED forms dn0 from dn slice 0 on device (bpm_3d.py:149-159)
 *  n0,k0,dx,dy,dz  float32 kernel arguments of the on-the-fly propagator
 *  unit_k     float32(k0*dz) of the phase screen
 *  absorb     taper width in pixels (0: none)
 *  out_field  (batch, nz, ny, nx) c64 or NULL     -- return_full, field
 *  out_intens (batch, nz, ny, nx) f32 or NULL     -- return_full, intensity
 *  out_numel  elements each non-NULL output volume holds (>= batch*nz*ny*nx)
 *  out_sum    (batch, ny, nx) f32 or NULL: sum over all nz planes of |u|^2
 *  scat_w,g_w (ny, nx) f32 cos(theta), cos(theta)^2; plain (nz) c64 the
 *             Nx*Ny*exp(-i k0 n0 z dz) DC term; out_scat (batch, nz, 2) f32
 *             receives the unnormalised sums (bpm_3d.py:227-260,335-341);
 *             all NULL when scattering is not requested.
 * The plan's scratch (weighted-mean and scattering partials) is allocated on
 * first use and grown when a call needs more; a plan serves one stream at a
 * time (calls on several streams need a plan each). */
int bpmh_propagate(bpmh_plan plan, int batch, int nz, float* plane, const void* dn,
                   int dn_kind, long long dn_batch_stride, long long dn_numel, int mean_mode,
                   const float* H, const float* n_eff, const float* dn0, float n0, float k0,
                   float dx, float dy, float dz, float unit_k, int absorb, float* out_field,
                   float* out_intens, long long out_numel, float* out_sum, const float* scat_w,
                   const float* g_w, const float* plain, float* out_scat, void* stream);

/* bpmh_propagate without any allocation: the same arguments and kernels
 * (bit-identical results), eager launches only (graph mode and timing are
 * ignored), and the step scratch -- weighted-mean and scattering partials --
 * taken from a fixed 16 MiB device array that each GPU holds for the life of
 * the process instead of plan memory.  A call whose scratch would exceed it
 * (it needs 8*batch*nx bytes for the weighted mean, 8*batch*nz*nx for
 * scattering sums, plus 256) returns BPMH_ERR_INVALID.  The plan's tables
 * must already be on the device (plans are created there when a device is
 * present).  Nothing in the array needs clearing between calls; calls that
 * overlap on one GPU (several streams or threads) share the array and race.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). */
int bpmh_propagate_noalloc(bpmh_plan plan, int batch, int nz, float* plane, const void* dn,
                           int dn_kind, long long dn_batch_stride, long long dn_numel,
                           int mean_mode, const float* H, const float* n_eff, const float* dn0,
                           float n0, float k0, float dx, float dy, float dz, float unit_k,
                           int absorb, float* out_field, float* out_intens, long long out_numel,
                           float* out_sum, const float* scat_w, const float* g_w,
                           const float* plain, float* out_scat, void* stream);

/* Ragged batch on the static-propagator (Bpm3d / mean_mode ignore) path:
 * n_groups (1-16) groups, group g holding items[g] fields that propagate
 * through their own dn volume, nzs[g] planes at dn + g*dn_group_stride
 * elements, with the groups in order of non-increasing depth.  The fields
 * are consecutive in plane (sum(items), ny, nx) c64 (u0 in, each field's
 * last plane out) and out_sum (sum(items), ny, nx) f32 or NULL (sum of
 * |u|^2 over the group's planes).  Every step runs only the items still
 * propagating (a prefix of them), so jobs of one grid size but different
 * depths and volumes share one launch sequence; each item gives the bits of
 * bpmh_propagate on its own volume (same plan variant).  group_done: NULL
 * or n_groups HIP events (hipEvent_t) to record on `stream` as soon as that
 * group's planes and sums are final.  dn_numel: elements dn holds (every
 * group's planes must lie inside).  No reference counterpart (the reference
 * propagates one field per call, bpm/bpm/bpm_3d.py:39-101); it serves
 * several requests of one grid size in one launch sequence. */
int bpmh_propagate_groups(bpmh_plan plan, int n_groups, const int* items, const int* nzs,
                          float* plane, const void* dn, int dn_kind, long long dn_group_stride,
                          long long dn_numel, const float* H, float unit_k, int absorb,
                          float* out_sum, void** group_done, void* stream);

/* Coated-sphere refractive-index volume (nz, ny, nx) float32: for each voxel
 * at (i dx, j dy, k dz), dmin = distance to the nearest of the n_points
 * centres (device float32 array n x 3, xyz, sorted by z) whose z is within
 * rad2; dn = dmin < rad1 ? dn1 : dmin < rad2 ? dn2 : 0.  Replaces the fill_dn
 * kernel (bpm/bpm/kernels/bpm_3d_spheres.cl:31-74) launched by
 * create_dn_buffer (bpm/bpm/bpm_3d_spheres.py:11-38). */
int bpmh_fill_dn_coated(int nx, int ny, int nz, float dx, float dy, float dz,
                        const float* points, int n_points, float dn1, float rad1, float dn2,
                        float rad2, float* dn, void* stream);

/* Summed-sphere volume: dn = sum_s dn_s [|r - c_s| < r_s] with spheres a
 * device float64 array n x 5 (cx, cy, cz, r, dn) in micrometres; distances
 * in float64, float32 sum in sphere order.  No reference counterpart (the
 * reference builds such volumes on the host); it generates the benchmark
 * workloads' dn on the device instead of uploading 4 GiB. */
int bpmh_spheres_dn(int nx, int ny, int nz, double dx, double dy, double dz,
                    const double* spheres, int n_spheres, float* dn, void* stream);

/* Adjoint split-step over planes z_lo..z_hi of a static-propagator
 * (mean_mode ignore / Bpm3d) run with dn shared by `batch` fields, dn_kind
 * BPMH_DN_F32 (real) or BPMH_DN_C64 (absorbing; no taper, as in the forward
 * pass).  g: (batch, ny, nx) c64, on entry G = 2 dL/du*[z_hi] (the gradient
 * convention of torch), on exit G for u[z_lo].  u_seg: the forward planes
 * u[z_lo..z_hi] (batch, z_hi-z_lo+1, ny, nx) c64 (recomputed from a
 * checkpoint by bpmh_propagate with out_field), u_seg_numel elements.  dn:
 * the (nz, ny, nx) volume; grad_dn, of dn's type, receives += dL/ddn for
 * planes z_lo+1..z_hi (complex dn: dL/dRe + i dL/dIm); both hold dn_numel
 * elements, at least (z_hi+1)*ny*nx.  gitem: (batch, ny, nx) scratch of dn's
 * type.  H: the forward table.
 * No reference counterpart (the reference has no gradient; its
 * bpm/bpm/bpm_3d_inverse.py:83-104 is a non-differentiable phase-ratio
 * experiment on absorbing dn). */
int bpmh_adjoint_segment(bpmh_plan plan, int batch, int z_lo, int z_hi, float* g,
                         const float* u_seg, long long u_seg_numel, const void* dn, int dn_kind,
                         long long dn_numel, const float* H, float unit_k, int absorb,
                         float* grad_dn, float* gitem, void* stream);

/* The differentiable form under bpm_3d's mean corrections, one segment
 * z_lo..z_hi of a float32 dn volume (nz, ny, nx) shared by the batch
 * (dn_numel elements, at least (z_hi+1)*ny*nx); n0, k0, dx, dy, dz, unit_k,
 * absorb as for bpmh_propagate.
 * mean_mode BPMH_MEAN_WEIGHTED (bpm_3d's default, bpm/bpm/bpm_3d.py:149-159,
 *   302-311): dn0 = sum|U| d / sum|U| over the k-space plane U = F u of each
 *   step, with d the volume's slice 0 read at the same flat index;
 * mean_mode BPMH_MEAN_SIMPLE (bpm_3d.py:216-217,288-293): dn0 = mean(dn[z])
 *   of the step's target plane, given as the float32 arrays n_eff (n0 + dn0)
 *   and dn0 over all nz planes, as for bpmh_propagate.
 * Each step applies H(n0 + dn0) and the screen exp(-i k0 dz (dn - dn0)).
 * g == NULL: forward -- plane (batch, ny, nx) c64 is advanced in place from
 *   plane z_lo to plane z_hi, bit-identical to those steps of bpmh_propagate
 *   in the same mean mode over the whole volume.
 * g != NULL: adjoint -- plane holds the forward plane z_lo and is only read;
 *   the segment is recomputed into work (keeping each step's U and, when
 *   weighted, its dn0 and sum|U|), then g (batch, ny, nx) c64 is carried from
 *   G_{z_hi} = 2 dL/du* of plane z_hi to G_{z_lo}, and grad_dn (nz, ny, nx)
 *   f32 receives += dL/ddn for planes z_lo+1..z_hi (through the screens and,
 *   simple, through each plane's mean) and, weighted, for slice 0 (through
 *   the mean's weights).
 * work: device floats, at least what bpmh_mean_segment_work reports
 * (forward: 2*batch*nx + batch; adjoint: about batch*ny*nx*(4*nseg + 4)).
 * Nothing is allocated; calls sharing a work buffer must not overlap.
 * No reference counterpart (the reference has no gradient). */
int bpmh_mean_segment(bpmh_plan plan, int mean_mode, int batch, int z_lo, int z_hi, float* plane,
                      const float* dn, long long dn_numel, const float* n_eff, const float* dn0,
                      float n0, float k0, float dx, float dy, float dz, float unit_k, int absorb,
                      float* g, float* grad_dn, float* work, long long work_numel, void* stream);
int bpmh_mean_segment_work(bpmh_plan plan, int batch, int z_lo, int z_hi, int adjoint,
                           long long* numel);

/* Debye-Wolf kernel variants */
#define BPMH_DW_INDEX 0  /* debye_wolf, bpm/psf/kernels/psf_debye.cl:24-113 */
#define BPMH_DW_SIN 1    /* debye_wolf_plane / debye_wolf_at, psf_debye.cl:117-269 */
#define BPMH_DW_GAUSS 2  /* debye_wolf_gauss, psf_debye.cl:273-362 */

/* Vectorial Debye-Wolf field on a regular (nz, ny, nx) grid with coordinates
 * x = x1 + i*(x2-x1)/(nx-1) (same for y, z).  Replaces the debye_wolf /
 * debye_wolf_gauss launches of bpm/psf/_focus_fields_debye.py:102-109,
 * :265-274, :404-413 (and debye_wolf_plane at :201-208 with nz=1, z1=z2).
 * alphas: HOST array of n_alphas float32 aperture angles (annulus pairs);
 * int_steps: the reference's compile-time INT_STEPS (runtime here).
 * Outputs ex, ey, ez (c64) and I (f32) may be NULL individually.
 * i_mode 0: I = |Ex|^2+|Ey|^2+|Ez|^2; 1: I = |Ex|^2 (debye_wolf_at). */
int bpmh_debye_wolf(int nx, int ny, int nz, float x1, float x2, float y1, float y2, float z1,
                    float z2, float lam, const float* alphas, int n_alphas, int int_steps,
                    float ex0, float ey0, int variant, float sig, int i_mode, float* ex,
                    float* ey, float* ez, float* I, void* stream);

/* focus_field_debye in one call: the debye_wolf quadrature on the octant
 * (nz0/2+1, ny0/2+1, nx0/2+1) at x = 0 .. (nx0/2)*dx (likewise y, z), spread
 * to the full (nz0, ny0, nx0) grid with the index map of bpmh_debye_mirror
 * (f -> |f - N0/2|, fields conjugated for f_z <= nz0/2) -- the same bits as
 * bpmh_debye_wolf on the octant followed by bpmh_debye_mirror, with the
 * octant in the library's stream-ordered scratch (the pool of
 * bpmh_debye_wolf's tables).  Replaces the debye_wolf launch and the numpy mirror of
 * bpm/psf/_focus_fields_debye.py:100-143 (x-polarised, Ex0 = 1, Ey0 = 0; lam
 * is the wavelength in the medium, lam/n0).  Even sizes only; outputs (full
 * grid, any may be NULL): ex, ey, ez c64, I f32. */
int bpmh_focus_field_debye(int nx0, int ny0, int nz0, double dx, double dy, double dz,
                           float lam, const float* alphas, int n_alphas, int int_steps,
                           float* ex, float* ey, float* ez, float* I, void* stream);

/* Measurement hooks of the grid path (bpmh_debye_wolf,
 * bpmh_focus_field_debye) on the calling thread: with timing on, every
 * kernel is bracketed by a HIP event pair; bpmh_debye_kernel_times waits for
 * them and returns the summed milliseconds and launch counts per kernel
 * kind since the last read-out in ms[5], counts[5] (0 tables: the sample
 * parameters, amplitudes, phases and azimuth terms in one launch, 1 product,
 * 2 full-grid spread of bpmh_focus_field_debye; 3 and 4 stay zero).  The
 * reference only
 * prints wall-clock deltas (bpm/psf/_focus_fields_debye.py:100,199,210). */
int bpmh_debye_set_timing(int enable);
int bpmh_debye_kernel_times(float* ms, int* counts);

/* Same integrand at n arbitrary points (device arrays x, y, z).  Replaces
 * debye_wolf_at, bpm/psf/_focus_fields_debye.py:346-352. */
int bpmh_debye_wolf_points(int n, const float* x, const float* y, const float* z, float lam,
                           const float* alphas, int n_alphas, int int_steps, float ex0,
                           float ey0, int variant, float sig, int i_mode, float* ex,
                           float* ey, float* ez, float* I, void* stream);

/* Spread an (nz0/2+1, ny0/2+1, nx0/2+1) octant to the full (nz0, ny0, nx0)
 * grid: index f -> |f - n0/2|, fields conjugated for f_z <= nz0/2.  Replaces
 * the numpy mirror of bpm/psf/_focus_fields_debye.py:116-143.  Any of the
 * four array pairs may be NULL. */
int bpmh_debye_mirror(int nx0, int ny0, int nz0, const float* u_oct, const float* ex_oct,
                      const float* ey_oct, const float* ez_oct, float* u, float* ex,
                      float* ey, float* ez, void* stream);

/* Cylindrical-lens focus field on an (nz, ny) grid (y = y1 + i (y2-y1)/(ny-1),
 * same for z; nz = 1 means z = z1): Ex = sum_t w sqrt(cos t) e^{i(ky sin t +
 * kz cos t)}, Ey, Ez the same times cos t, sin t, I = |Ex|^2+|Ey|^2+|Ez|^2,
 * S+1 samples over [-alpha, alpha + dt].  plane_variant 0 replaces the
 * psf_cylindrical launch (bpm/psf/_focus_fields_cylindrical.py:61-68), 1 the
 * psf_cylindrical_plane launch (:124-130) whose end weights test t == -+alpha.
 * Any output may be NULL. */
int bpmh_cylindrical(int ny, int nz, float y1, float y2, float z1, float z2, float lam,
                     float alpha, int int_steps, int plane_variant, float* ex, float* ey,
                     float* ez, float* I, void* stream);

/* out[i] = J_n(x[i]) for n in {0,1,2} with this library's float32 Bessel
 * evaluation, the one its Debye tables use (device arrays).  Replaces the bessel_fill self-test kernel,
 * bpm/psf/kernels/bessel.cl:120-129 (bpm/psf/_focus_fields_debye.py:34-42). */
int bpmh_bessel_fill(int n, const float* x, float* out, int count, void* stream);

/* Host (CPU) evaluation of the same Bessel code, for validating it without a
 * GPU (double precision in and out; host arrays).  Not part of any hot path. */
int bpmh_bessel_host(int n, const double* x, double* out, int count);

#ifdef __cplusplus
}
#endif
#endif /* BPM_HIP_H */
