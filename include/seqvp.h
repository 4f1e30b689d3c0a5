/* C-ABI of libseqvp_hip.so: the MI355X (gfx950) kernels of the seq_vp_3d registration hot path.
 *
 * Conventions (every entry point):
 *   - plain device pointers and sizes, no torch types; the caller owns every buffer (kernels never
 *     allocate or free); `stream` is a hipStream_t passed as void* (NULL = the null stream);
 *   - node fields are (nz, ny, nx, 3) float64, components xyz, x fastest (jd_curl layout);
 *     free-node vectors are component-major (3, nz-2, ny-2, nx-2);
 *   - `k` points to the 39-double Q1 constant vector in device memory:
 *     k[0:3]=1/h (xyz), k[3:6]=Gauss nodes, k[6:33]=Gauss weights omega*wx*wy*wz (point 9iz+3iy+ix),
 *     k[33]=mu, k[34]=nu, k[35]=barrier weight kappa*omega/27, k[36]=delta, k[37]=activation,
 *     k[38]=activation-delta;
 *   - return 0 on success, a negative SEQVP_E* code otherwise; nothing throws across the ABI;
 *   - all work is enqueued on `stream`; results are valid after that stream is synchronised.
 *   - the library holds no mutable global state, so calls are reentrant across host threads/streams; the one
 *     exception is seqvp_q1_soft_energy_noalloc's fixed scratch array (see there).
 */
#ifndef SEQVP_H
#define SEQVP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SEQVP_OK 0
#define SEQVP_EINVAL -1     /* bad size/shape/pointer */
#define SEQVP_ELAUNCH -2    /* kernel launch failed */
#define SEQVP_EWORKSPACE -3 /* workspace too small */

/* Version string of the library build. */
const char* seqvp_version(void);

/* Workspace bytes an entry point needs for a grid of nx*ny*nz nodes.
 * op: 0 soft energy, 1 barrier energy, 2 gradient, 3 GGN diagonal, 4 joint outer gradients,
 *     5 JD root, 6 DST apply, 7 dot products, 8 node-min JD, 10 trial energy (soft + barrier). */
size_t seqvp_workspace_bytes(int op, int64_t nx, int64_t ny, int64_t nz);

/* Soft JD/curl change energies (replaces FusedQ1.soft_energy,
 * jd_curl/three_d/backends/triton_q1.py:572-609; formulas residuals.py:166-172 and
 * energy_comparison.py:196-200).  out[0] = 1/2 mu sum qw (J-Jr)^2, out[1] = 1/2 nu sum qw |c-cr|^2,
 * out[2] = sum 1/2 (y-x)(y+x) over the JD/curl residuals of u (y) and base (x); 0 when base == NULL.
 * Sums are compensated (double-double). `out` is 3 device doubles. */
int seqvp_q1_soft_energy(const double* u, const double* ref, const double* base, const double* k,
                         int64_t nx, int64_t ny, int64_t nz, double* out, void* workspace,
                         size_t workspace_bytes, void* stream);

/* The trial's soft and barrier energies in one sweep (replaces the pair FusedQ1.soft_energy + barrier_energy of
 * AcceleratedHybridBackend._energy, accelerated.py:1574-1664): out[0:3] = seqvp_q1_soft_energy's [jd, curl,
 * difference], out[3:7] = seqvp_q1_barrier_energy's [value, min J, unweighted change, min base J], bit for bit.
 * `out` is 7 device doubles; workspace op 10 (12 doubles per 256-cell block, 256 bytes, and 192 doubles for the
 * reduction's stage). */
int seqvp_q1_trial_energy(const double* u, const double* ref, const double* base, const double* k,
                          int64_t nx, int64_t ny, int64_t nz, double* out, void* workspace,
                          size_t workspace_bytes, void* stream);

/* seqvp_q1_trial_energy for a base whose least closed27 J is known (the same trial energy call,
 * triton_q1.py:572-609, given what the previous evaluation of that state returned as its min J): `base_min` is one
 * device double holding the base's least closed27 J over the grid (NULL: unknown). quiet != 0 states that the
 * caller has checked *base_min is at or above the barrier activation k[37], so that no base point is active: the
 * sweep then evaluates no base closed27 J and out[6] is *base_min (if *base_min is below the activation, out[3] is
 * NaN). Otherwise the sweep skips a base cell's closed-point J where the cell's edges bound it above both the
 * activation and *base_min. Every output is seqvp_q1_trial_energy's bit for bit. (Any call also skips the
 * closed-point J of trial cells whose edges bound it above the activation and above the least J the sweep has
 * already met.) quiet requires a base other than ref and a base_min. Workspace op 10. */
int seqvp_q1_trial_energy_base_min(const double* u, const double* ref, const double* base, const double* base_min,
                                   int quiet, const double* k, int64_t nx, int64_t ny, int64_t nz, double* out,
                                   void* workspace, size_t workspace_bytes, void* stream);

/* One pair of a batched trial-energy launch: the arguments of seqvp_q1_trial_energy (base required). */
typedef struct {
  const double* u;
  const double* ref;
  const double* base;
  const double* k;
  int64_t nx, ny, nz;
  double* out;               /* 7 device doubles */
  void* workspace;           /* the pair's own workspace, op 10 */
  size_t workspace_bytes;
} SeqvpTrialBatchEntry;

/* The trial energies of 1..32 pairs of any sizes in one persistent launch (the batching form; no counterpart in
 * the reference, which evaluates one pair per call, triton_q1.py:572-609): the work list is every pair's 256-cell
 * blocks, `tile` (1..64) consecutive blocks per work item, taken by `grid` workgroups (rounded up to a multiple of 8;
 * 0 = one workgroup per item), then one reduction launch for all pairs. Each pair's out equals its
 * seqvp_q1_trial_energy result bit for bit, whatever tile and grid. */
int seqvp_q1_trial_energy_batch(const SeqvpTrialBatchEntry* pairs, int npairs, int tile, int64_t grid, void* stream);

/* seqvp_q1_soft_energy without a workspace and without any allocation. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). The block partials live in a fixed 4 MiB device array of the library (one per GPU, resolved on
 * the first call on that device and held until the process ends), so grids are limited to (nx-1)(ny-1)(nz-1) <=
 * 22,369,536 cells (87,381 blocks of 256 cells); a larger grid returns SEQVP_EWORKSPACE and launches nothing.
 * Same kernels, same results (bit for bit) as seqvp_q1_soft_energy. Calls that overlap on one GPU (several
 * streams or host threads at once) share that array and race; serialise them. */
int seqvp_q1_soft_energy_noalloc(const double* u, const double* ref, const double* base, const double* k,
                                 int64_t nx, int64_t ny, int64_t nz, double* out, void* stream);

/* Closed27 localized barrier energy (replaces FusedQ1.barrier_energy, triton_q1.py:680-719;
 * barrier.py:19-36, energy_comparison.py:14-46).  out[0] = w sum psi(J), out[1] = min J,
 * out[2] = UNWEIGHTED sum of stable psi(J)-psi(Ja) vs base (the caller multiplies by w, as
 * accelerated.py:1627 does), out[3] = min base J (0 / +inf when base == NULL).
 * A NaN J propagates to out[1]. `out` is 4 device doubles. */
int seqvp_q1_barrier_energy(const double* u, const double* base, const double* k, int64_t nx, int64_t ny,
                            int64_t nz, double* out, void* workspace, size_t workspace_bytes, void* stream);

/* FP32 screening variants of the two energies above (replace FusedQ1.soft_energy_f32 / barrier_energy_f32,
 * triton_q1.py:652-678; accelerated.py:1603-1612): node values rounded to binary32 on load, per-point
 * terms in binary32, per-lane Kahan-compensated binary32 sums, block/grid sums in double-double.
 * Same outputs and workspace as the FP64 entry points. */
int seqvp_q1_soft_energy_f32(const double* u, const double* ref, const double* base, const double* k, int64_t nx,
                             int64_t ny, int64_t nz, double* out, void* workspace, size_t workspace_bytes,
                             void* stream);
int seqvp_q1_barrier_energy_f32(const double* u, const double* base, const double* k, int64_t nx, int64_t ny,
                                int64_t nz, double* out, void* workspace, size_t workspace_bytes, void* stream);

/* Node-term energies (replaces the image/proximity part of AcceleratedHybridBackend._energy,
 * accelerated.py:1583-1600; residuals.py:189-195, energy_comparison.py:196-210).
 * q is the trapezoid node mass omega*tz*ty*tx computed in-kernel; qd = q*data_weight (data_weight may
 * be NULL = 1).  out[0] = 1/2 sum qd (W-F)^2, out[1] = 1/2 lam sum q |u-ref|^2, and with a base
 * (base_warped and base_u non-NULL): out[2] = sum 1/2 (y-x)(y+x) of the image residuals (outer F
 * block), out[3] = the same over the proximity residuals. `out` is 4 device doubles. */
int seqvp_node_energy(const double* warped, const double* fixed, const double* data_weight, const double* u,
                      const double* ref, const double* base_warped, const double* base_u, double lam,
                      double hx, double hy, double hz, int64_t nx, int64_t ny, int64_t nz, double* out,
                      void* workspace, size_t workspace_bytes, void* stream);

/* Node terms of the gradient and the GGN diagonal (replaces the node-term lines of
 * TorchLinearization.gradient / joint_outer_gradients / diagonal, torch_ops.py:563-568, 591-600, 784-791;
 * linearization.py:150-156, 362): image = qd (W-F) b, total = image + lam q (u-ref), ggn = qd b^2 + lam q,
 * with q the trapezoid node mass and qd = q*data_weight (data_weight NULL = 1). Outputs may be NULL. */
int seqvp_node_terms(const double* warped, const double* fixed, const double* data_weight, const double* image_grad,
                     const double* u, const double* ref, double lam, double hx, double hy, double hz, int64_t nx,
                     int64_t ny, int64_t nz, double* image, double* total, double* ggn, void* stream);

/* Energy gradient on the nodes (replaces FusedQ1.gradient_nodes, triton_q1.py:504-518;
 * linearization.py:159-201): out = node_base + [Gauss JD/curl VJP] + [closed27 barrier VJP].
 * node_base may be NULL (zero). out may alias node_base. out is (nz,ny,nx,3). */
int seqvp_q1_gradient(const double* node_base, const double* u, const double* ref, int soft, int barrier,
                      const double* k, int64_t nx, int64_t ny, int64_t nz, double* out, void* workspace,
                      size_t workspace_bytes, void* stream);

/* The energy gradient on the free nodes in one call (replaces TorchLinearization.gradient followed by the
 * free-node restriction, torch_ops.py:559-568; linearization.py:150-201): g = restrict(node terms total + Gauss
 * JD/curl VJP + closed27 barrier VJP), g component-major (3, nz-2, ny-2, nx-2). The node terms are those of
 * seqvp_node_terms and the gather order that of seqvp_q1_gradient; the node field is never written. */
int seqvp_q1_gradient_free(const double* warped, const double* fixed, const double* data_weight,
                           const double* image_grad, const double* u, const double* ref, double lam, double hx,
                           double hy, double hz, const double* k, int64_t nx, int64_t ny, int64_t nz, double* g,
                           void* workspace, size_t workspace_bytes, void* stream);

/* The outer gradients on the free nodes in one call (replaces the node terms, FusedQ1.joint and the three
 * restrictions of the outer norms, linearization.py:217-238; sequential.py:288-338): gd = restrict(image term),
 * gf = restrict(image term + barrier VJP), ge = restrict(total term + Gauss JD/curl VJP + barrier VJP), each
 * component-major (3, nz-2, ny-2, nx-2), the gather order that of seqvp_q1_joint. */
int seqvp_q1_joint_free(const double* warped, const double* fixed, const double* data_weight, const double* image_grad,
                        const double* u, const double* ref, double lam, double hx, double hy, double hz,
                        const double* k, int64_t nx, int64_t ny, int64_t nz, double* gd, double* gf, double* ge,
                        void* workspace, size_t workspace_bytes, void* stream);

/* GF and GE node gradients sharing the cell work (replaces FusedQ1.joint, triton_q1.py:563-570;
 * linearization.py:217-238): outer = outer_base + barrier VJP; total = total_base + soft VJP +
 * barrier VJP. */
int seqvp_q1_joint(const double* outer_base, const double* total_base, const double* u, const double* ref,
                   const double* k, int64_t nx, int64_t ny, int64_t nz, double* outer, double* total,
                   void* workspace, size_t workspace_bytes, void* stream);

/* Exact GGN diagonal on the nodes (replaces FusedQ1.diagonal_nodes, triton_q1.py:442-458;
 * linearization.py:357-380): out = node_base + per-cell Gauss soft + closed27 barrier diagonal. */
int seqvp_q1_ggn_diagonal(const double* node_base, const double* u, const double* k, int64_t nx, int64_t ny,
                          int64_t nz, double* out, void* workspace, size_t workspace_bytes, void* stream);

/* T62 node JD measure without the T69 extension (replaces the closed27/edge part of
 * TorchLinearization.node_min_closed27_jd, torch_ops.py:687-773; jd_affine.py:24-48,205-225):
 * node[n] = min over the cells of n of min(closed27 J, exact edge-quadratic minima); +inf if none. */
int seqvp_node_min_jd(const double* u, const double* k, int64_t nx, int64_t ny, int64_t nz, double* node,
                      void* workspace, size_t workspace_bytes, void* stream);

/* The same measure with the T69 face/interior extension (replaces jd_affine.node_min_jd with
 * extension_activation, jd_affine.py:154-225, and the backend's node_min_jd(lin, extension_activation=...,
 * stats=...), hybrid.py:97-101): on every cell whose Bernstein hull reaches below min(m_c, activation),
 * m_c <- min(m_c, mu_c) with mu_c the face/interior estimate of seqvp_face_interior_minima.
 * activation in (0, 1]. stats (device, 4 doubles, may be NULL): number of candidate cells, cells lowered by
 * more than 0.1 %, cells lowered at all, least mu (+inf without candidates). Workspace:
 * seqvp_node_min_jd_extended_workspace_bytes. */
size_t seqvp_node_min_jd_extended_workspace_bytes(int64_t nx, int64_t ny, int64_t nz);
int seqvp_node_min_jd_extended(const double* u, const double* k, int64_t nx, int64_t ny, int64_t nz,
                               double activation, double* node, double* stats, void* workspace,
                               size_t workspace_bytes, void* stream);

/* T69 face/interior estimate per cell (replaces jd_affine.face_interior_minima, jd_affine.py:130-151):
 * J27 holds n rows of 27 closed-point Jacobians (index 9 iz + 3 iy + ix); mu[i] = least of the row's
 * triquadratic on the 9^3 grid and of bounded Newton from its 4 lowest grid points. Device pointers. */
int seqvp_face_interior_minima(const double* J27, int64_t n, double* mu, void* stream);
/* HOST: the same estimate of host rows on the calling thread (the reference's CPU estimator,
 * jd_affine.py:163-172). */
int seqvp_face_interior_minima_host(const double* J27, int64_t n, double* mu);

/* JD witness nomination (replaces the tile loop of witness_screen.screen, witness_screen.py:55-76): for each tile
 * of `tile` consecutive cells, out[26 t] = least closed27 J of the tile (nonfinite J counts as +inf),
 * out[26 t + 1] = first flat index (cell - tile start) * 27 + point where it occurs, out[26 t + 2 ...] = the 24
 * vertex values (vertex-major, x + 2y + 4z) of that cell. out holds 26 * ceil(ncell / tile) doubles. */
int seqvp_closed27_tile_minima(const double* u, const double* k, int64_t nx, int64_t ny, int64_t nz, int64_t tile,
                               double* out, void* stream);

/* Cubic coefficients of det(A + s V) - delta at n reference points (x, y, z) of cells (replaces
 * determinant_coefficients over JDStepMemory.predict's witness cells, jd_step.py:13-17, 204-222): A = I + grad u,
 * V = grad v (node fields), delta = k[36]. out: (n, 4) ascending powers; NaN for a cell id outside the grid. */
int seqvp_witness_coefficients(const double* u, const double* v, const double* k, int64_t nx, int64_t ny, int64_t nz,
                               const int64_t* cells, const double* points, int64_t n, double* out, void* stream);

/* T66 global closed27 first JD root (replaces FusedJDRoot.__call__, triton_jd_root.py:38-52;
 * global_jd_step.py:22-129). u: base field, v: direction field (both node fields), limit > 0.
 * out[0] = min root in s-units (inf if none), out[1] = number of points with c0 <= 0 (as double),
 * out[2] = 1.0 if any cubic coefficient was not finite, else 0.0. `out` is 3 device doubles. */
int seqvp_jd_root(const double* u, const double* v, const double* k, double limit, int64_t nx, int64_t ny,
                  int64_t nz, double* out, void* workspace, size_t workspace_bytes, void* stream);

/* seqvp_jd_root with the direction given by its free-node vector p (component-major, 3 x (nx-2)(ny-2)(nz-2)):
 * the same root, bit for bit, as seqvp_jd_root on inject(p), without writing the node field. */
int seqvp_jd_root_free(const double* u, const double* p, const double* k, double limit, int64_t nx, int64_t ny,
                       int64_t nz, double* out, void* workspace, size_t workspace_bytes, void* stream);

/* seqvp_jd_root_free with the limit read from device memory (`limit`: one double, written by seqvp_pg_step), so the
 * sweep can follow the direction without the host reading the step first. A limit that is not positive and finite
 * gives results the caller must discard (the host's driver stops before such a sweep). Same kernels, same bits. */
int seqvp_jd_root_free_dev(const double* u, const double* p, const double* k, const double* limit, int64_t nx,
                           int64_t ny, int64_t nz, double* out, void* workspace, size_t workspace_bytes,
                           void* stream);

/* T75 Bernstein floating-point prescreen (replaces prescreen_lower_bounds,
 * jd_curl/three_d/certificate_prescreen.py:125-153; kernel :60-121), bitwise in IEEE binary64 with no
 * contraction and the reference operation order. D: closed27 derivative table (3,27,8) (device).
 * cells: increasing int64 cell ids (device) or NULL for all cells (then ncells must equal the cell
 * count). out[i] = min beta_fl - B for cell i, NaN where a value is not finite. */
int seqvp_prescreen(const double* u, const double* D, const int64_t* cells, int64_t ncells, int64_t nx,
                    int64_t ny, int64_t nz, double* out, void* stream);

/* Prescreen of all cells fused with the proof test: flags[c] = 1 if min beta_fl - B > delta (proven),
 * else 0; out_count[0] = number of unproven cells (int64, device). Same arithmetic as seqvp_prescreen. */
int seqvp_prescreen_flags(const double* u, const double* D, double delta, int64_t nx, int64_t ny, int64_t nz,
                          uint8_t* flags, int64_t* out_count, void* stream);

/* Tensor cubic B-spline value and gradient at node positions X + u (replaces TorchSpline.sample /
 * sample_points, jd_curl/three_d/backends/spline.py:81-127; interpolation.py:72-113).
 * coeff: (cz, cy, cx) coefficients; tz/ty/tx: knot vectors of lengths ntz/nty/ntx (cubic, so
 * coefficient counts are nt-4). origin/spacing: xyz of the node grid (nx,ny,nz) whose positions
 * are sampled. value: (N) or NULL; grad: (N,3) xyz or NULL. status[0] (int64, device) counts queries
 * outside the closed support, or, for free nodes, outside the open support (0 = all valid). */
int seqvp_spline_warp(const double* u, const double* coeff, int64_t cz, int64_t cy, int64_t cx,
                      const double* tz, int64_t ntz, const double* ty, int64_t nty, const double* tx, int64_t ntx,
                      double ox, double oy, double oz, double hx, double hy, double hz, int64_t nx, int64_t ny,
                      int64_t nz, double* value, double* grad, int64_t* status, void* stream);

/* Same sampler at arbitrary points (npts, 3) xyz (replaces TorchSpline.sample, backends/spline.py:81-127); status[0]
 * counts points outside the closed box [lox,hix]x[loy,hiy]x[loz,hiz], and, where `strict` (npts bytes, NULL = none)
 * is nonzero, outside the open box; their value/gradient are NaN. */
int seqvp_spline_points(const double* points, int64_t npts, const uint8_t* strict, const double* coeff, int64_t cz,
                        int64_t cy,
                        int64_t cx, const double* tz, int64_t ntz, const double* ty, int64_t nty,
                        const double* tx, int64_t ntx, double lox, double hix, double loy, double hiy,
                        double loz, double hiz, double* value, double* grad, int64_t* status, void* stream);

/* DST-I Poisson inverse (replaces TorchPoisson.apply, torch_ops.py:107-115 and
 * _HostPoisson.apply_equilibrated, hybrid.py:262-286; PoissonInverse3D.apply, poisson.py:99-106):
 *   out = sign * s * IDST( DST(s * rhs) / den ),  DST = orthonormal type-I along x, y, z,
 * rhs/out: (3, mz, my, mx) (mz,my,mx = interior sizes), den: (mz, my, mx), s: same layout as rhs or
 * NULL (= 1), sign = -1 when `negate` else +1. Dense sine-matrix contractions on MFMA (FP64).
 * out may alias rhs. */
int seqvp_dst_poisson(const double* rhs, const double* den, const double* scale, int negate, int64_t mx,
                      int64_t my, int64_t mz, int64_t ncomp, double* out, void* workspace,
                      size_t workspace_bytes, void* stream);

/* Spectral denominator omega*(a + b*lambda) of the Q1 (q1 = 1) or seven-point (q1 = 0) metric on the
 * interior DST-I modes, built on the device from its separable 1-D factors lx/wx (length mx), ly/wy (my),
 * lz/wz (mz) (replaces poisson_spectrum and the denominator of build_poisson_inverse, poisson.py:16-30,
 * 118-131; the products are formed in the reference's order without contraction). den: (mz, my, mx). */
int seqvp_poisson_denominator(const double* lx, const double* wx, const double* ly, const double* wy,
                              const double* lz, const double* wz, int64_t mx, int64_t my, int64_t mz, int q1,
                              double omega, double a, double b, double* den, void* stream);

/* FP32 variant for the precision form (replaces TorchPoisson.apply_fp32, torch_ops.py:117-136):
 * rhs/out float64 in memory, transforms and division in float32 on MFMA. */
int seqvp_dst_poisson_f32(const double* rhs, const double* den, const double* scale, int negate, int64_t mx,
                          int64_t my, int64_t mz, int64_t ncomp, double* out, void* workspace,
                          size_t workspace_bytes, void* stream);

/* Compensated dot products of n-vectors: out[i] = sum x_i * y_i for the pairs given (npairs <= 4);
 * xs/ys are arrays of device pointers in host memory. */
int seqvp_dots(const double* const* xs, const double* const* ys, int npairs, int64_t n, double* out,
               void* workspace, size_t workspace_bytes, void* stream);

/* theta (3, nz-2, ny-2, nx-2) -> node field (nz, ny, nx, 3) with a zero boundary (boundary.py:8-12). */
int seqvp_inject(const double* theta, int64_t nx, int64_t ny, int64_t nz, double* u, void* stream);

/* node field -> theta on the free nodes (boundary.py:15-17; linearization.py:84-88). */
int seqvp_restrict(const double* u, int64_t nx, int64_t ny, int64_t nz, double* theta, void* stream);

/* Candidate map of the line search, u = inject(theta + tau p) with theta_out = theta + tau p when
 * theta_out != NULL; the product and sum are rounded separately, as the host axpy does
 * (poisson_gradient.py:1403-1437). */
int seqvp_axpy_inject(const double* theta, const double* p, double tau, int64_t nx, int64_t ny, int64_t nz,
                      double* theta_out, double* u, void* stream);

/* The scalars of a PG direction in one pass over the free vectors (the host products of poisson_gradient.py:955-1011
 * and the slope/gradient-norm reductions of :1263, 1307-1315): out[0] = g.p, out[1] = g.g; with the previous
 * accepted state (theta_old, g_old, p_old; all NULL for none) out[4..7] = seqvp_bb2_products' four values and
 * out[9..10] = seqvp_max_abs_diff(theta, theta_old)'s two; out[8] = seqvp_step_bound(p, u)'s bound. Every value is
 * the bits of its separate entry point; out has 11 device doubles (2, 3 untouched); workspace op 7. */
int seqvp_direction_scalars(const double* theta, const double* theta_old, const double* g, const double* g_old,
                            const double* p_old, const double* p, const double* u, double ox, double oy, double oz,
                            double hx, double hy, double hz, int64_t nx, int64_t ny, int64_t nz, double* out, void* ws,
                            size_t ws_bytes, void* stream);

/* The first attempt's step of a PG iteration on the device (the host step of poisson_gradient.py:955-990 and
 * 1340-1356): from `scalars` as seqvp_direction_scalars leaves them (BB2 products out[4..7], bound out[8]), the BB2
 * proposal with the reference's fallbacks (has_prev = 0: initial_step; else 2 min(prev_tau, step_max/2) when the
 * secant is nonfinite, has s.y <= 0, cosine <= 1e-12 or y.iy <= 0, or s.y / y.iy is not positive), clipped to
 * [step_min, step_max], then min(proposal, 0.9 bound) into tau (one device double). The operations are the host
 * solver's, in its order, so both give the same bits. */
int seqvp_pg_step(const double* scalars, int has_prev, double prev_tau, double step_min, double step_max,
                  double initial_step, double* tau, void* stream);

/* The step capped by the T66 root (poisson_gradient.py:1356-1366): tau = min(limit, safety * root_out[0] * limit)
 * when root_out (seqvp_jd_root_free's out) holds a finite nonzero root and no nonfinite flag; else tau = limit.
 * limit and tau: one device double each. */
int seqvp_capped_step(const double* limit, const double* root_out, double safety, double* tau, void* stream);

/* The same candidate map with the step read from device memory (tau: one double), so a captured trial segment
 * (seqvp_graph_*) takes a new step on every launch; same arithmetic, same bits. Counterpart of the host axpy of
 * poisson_gradient.py:1403-1437 inside the device trial of accelerated.py:1666-1730. */
int seqvp_axpy_inject_dev(const double* theta, const double* p, const double* tau, int64_t nx, int64_t ny, int64_t nz,
                          double* theta_out, double* u, void* stream);

/* BB2 secant products (replaces the host products of _step_proposal, poisson_gradient.py:955-990):
 * with s = theta - theta_old, y = g - g_old, iy = p_old - p (each rounded once), out[0] = s.y,
 * out[1] = s.s, out[2] = y.y, out[3] = y.iy, compensated (double-double) sums. */
int seqvp_bb2_products(const double* theta, const double* theta_old, const double* g, const double* g_old,
                       const double* p_old, const double* p, int64_t n, double* out, void* workspace,
                       size_t workspace_bytes, void* stream);

/* Exact linear node-domain step bound (replaces sampling_step_bound, poisson_gradient.py:993-1011):
 * out[0] = min over free nodes/components of (hi - x)/p (p > 0) and (lo - x)/p (p < 0), x = X + u;
 * +inf when p = 0. p is the free vector, u the node field, origin/spacing the grid's. */
int seqvp_step_bound(const double* p, const double* u, double ox, double oy, double oz, double hx, double hy,
                     double hz, int64_t nx, int64_t ny, int64_t nz, double* out, void* workspace,
                     size_t workspace_bytes, void* stream);

/* out[0] = max|a - b|, out[1] = max|b| over n entries (NaN propagates) -- the relative step of
 * step_ratio (solvers/common.py:63-73). */
int seqvp_max_abs_diff(const double* a, const double* b, int64_t n, double* out, void* workspace,
                       size_t workspace_bytes, void* stream);

/* PG metric equilibration scale on the free vector (replaces the scale lines of _pg_metric_inverse,
 * poisson_gradient.py:472-479, 507-515, and affine_factors, jd_affine.py:228-242):
 * scale = (sqrt(m)/sqrt(diag)) * r, r = min(1, max(J_a, floor)/act) with J_a the node JD measure
 * (node may be NULL: r = 1). diag/scale are (3, nz-2, ny-2, nx-2); bad[0] (int64, device) counts
 * free entries with diag <= 0 or nonfinite diag / J_a (the reference raises FloatingPointError). */
int seqvp_metric_scale(const double* diag, const double* node, double m, double act, double floor_, int64_t nx,
                       int64_t ny, int64_t nz, double* scale, int64_t* bad, void* stream);

/* One v_mfma_f64_16x16x4f64 (fp32 = 0) or v_mfma_f32_16x16x4f32 (fp32 = 1) on A (16x4, row-major a)
 * and B (4x16, row-major b), written to d (16x16, row-major) through the lane mapping the DST kernels
 * assume; a layout self-check. */
int seqvp_mfma_probe(const double* a, const double* b, int fp32, double* d, void* stream);

/* Separable linear map of a scalar node field (nz, ny, nx): out = (Mz (x) My (x) Mx) in, each axis line multiplied by
 * that axis's dense row-major matrix; mx_t is Mx transposed. With the inverse collocation matrices of the not-a-knot
 * cubic interpolant it computes the image spline coefficients of build_image_sampler (interpolation.py:29-40,
 * make_interp_spline along z, y, x) on the device (FP64 MFMA). ws: nx*ny*nz doubles; in != out. */
int seqvp_axis_transform(const double* in, double* out, const double* mx_t, const double* my, const double* mz,
                         int64_t nx, int64_t ny, int64_t nz, void* ws, size_t ws_bytes, void* stream);

/* ---- whole-cell JD certificate (replaces certify_map with PrescreenCertificateConfig,
 *      jd_curl/three_d/feasibility.py:395-769; compiled_enclosure.py; compiled_fallback.py) ---------------- */

/* Certificate reasons (FeasibilityResult.reason strings). */
#define SEQVP_CERT_POSITIVE_ENCLOSURE_COVER 0 /* certified; some cells proven by the hull or the prescreen */
#define SEQVP_CERT_POSITIVE_BERNSTEIN_COVER 1 /* certified by the Bernstein fallback alone */
#define SEQVP_CERT_JACOBIAN_VIOLATION 2       /* a cell holds a point with J <= delta */
#define SEQVP_CERT_ARITHMETIC_UNCERTAIN 3     /* an interval step was nonfinite or inverted */
#define SEQVP_CERT_SUBDIVISION_BUDGET 4       /* a cell exhausted its box or depth budget */
#define SEQVP_CERT_NONFINITE_STATE 5
#define SEQVP_CERT_NONZERO_BOUNDARY 6
#define SEQVP_CERT_INVALID_DOMAIN 7

/* Map certificate (the FeasibilityResult/PrescreenFeasibilityResult fields, feasibility.py:136-173).
 * status: 0 certified_feasible, 1 uncertain, 2 infeasible. NaN stands for None. */
typedef struct {
  int32_t status;
  int32_t reason;             /* SEQVP_CERT_* */
  int32_t max_depth;
  int32_t arithmetic_failure;
  int64_t n_cells_checked;
  int64_t n_uncertain;
  int64_t n_infeasible;
  int64_t n_boxes;
  int64_t n_fast_cells;       /* proven by the edge hull */
  int64_t n_prescreen_cells;  /* proven by the T75 prescreen */
  int64_t n_fallback_cells;   /* classified by the interval Bernstein fallback */
  int64_t witness_cell;       /* first infeasible cell in tile order, -1 when none */
  double witness_point[3];    /* its reference point (x, y, z) in [0,1]^3 */
  double min_margin;          /* min_margin_lower_bound: lower bound of min (J - delta) */
  double min_jd;              /* min_jd_lower_bound */
  double sampled_min_jd;      /* float64 sample minimum over fallback cells (NaN when any cell was proven early) */
} SeqvpCertificate;

/* One cell's Bernstein classification (CellCertificate, feasibility.py:124-133).
 * reason: 0 positive_bernstein_cover, 1 sample_witness, 2 subdivision_witness, 3 box_budget, 4 depth_budget,
 * 5 arithmetic_uncertain. */
typedef struct {
  int32_t status;
  int32_t reason;
  int32_t max_depth;
  int32_t has_witness;
  int64_t n_boxes;
  double witness_point[3];
  double min_margin;
  double min_jd;
  double sampled_min_jd;
} SeqvpCellCertificate;

/* Device workspace of seqvp_certificate_screen (tile = cells per certificate tile, 8192 in the reference). */
size_t seqvp_certificate_workspace_bytes(int64_t nx, int64_t ny, int64_t nz, int64_t tile);

/* Edge-hull margins on the device, bitwise the CPU kernel's (replaces parallel_margins,
 * compiled_enclosure.py:193-227, 336-345, and the GPU edge hull of backends/edge_hull_gpu.py):
 * margin[c] = lo(J_hull - delta), flag[c] = 1 when an interval step was nonfinite or inverted (flag may be NULL). */
int seqvp_edge_hull(const double* u, double hx, double hy, double hz, double delta, int64_t nx, int64_t ny,
                    int64_t nz, double* margin, uint8_t* flag, void* stream);

/* Device screen of a candidate map (the hull and prescreen stages of _certify_map_parallel,
 * feasibility.py:571-695): edge hull with the whole-tile rule, then the T75 prescreen (D = closed27 derivative
 * table (3,27,8) on the device, NULL = no prescreen) of the cells the hull left. summary (5 device doubles) =
 * [min proven hull margin, min proven prescreen margin (+inf: none), n_fast, n_prescreen, n_undecided];
 * undecided[0 : min(n_undecided, capacity)] receives the undecided cell ids in no particular order;
 * codes (optional, ncell bytes) = 0 hull, 1 prescreen, 2 undecided. */
int seqvp_certificate_screen(const double* u, const double* D, double hx, double hy, double hz, double delta,
                             int64_t nx, int64_t ny, int64_t nz, int64_t tile, uint8_t* codes, int64_t* undecided,
                             int64_t capacity, double* summary, void* workspace, size_t workspace_bytes, void* stream);

/* out (n, 8, 3) = the vertex displacements of cells[0:n] (vertex a = x + 2y + 4z), on the device
 * (the flat[_node_indices_for(grid, cells)] gather of feasibility.py:712-713). */
int seqvp_gather_vertices(const double* u, const int64_t* cells, int64_t n, int64_t nx, int64_t ny, int64_t nz,
                          double* out, void* stream);

/* HOST. The interval arrays of the Bernstein fallback for n cells of host vertices (n, 8, 3): p = J - delta at
 * the 27 points, its Bernstein coefficients and the float64 sample J, each (n, 27) (replaces
 * compiled_fallback.fallback_arrays / feasibility._fallback_arrays, feasibility.py:515-523).
 * Returns 1 when every step of every cell was finite and ordered, 0 otherwise, SEQVP_EINVAL on bad arguments. */
int seqvp_bernstein_arrays(const double* vertices, int64_t n, const double* spacing, double delta, double* p_lo,
                           double* p_hi, double* c_lo, double* c_hi, double* samples);

/* HOST. Outward enclosure [out[0], out[1]] of J at the reference point (x, y, z) in [0,1]^3 of a cell with host
 * vertices (8, 3) (replaces stationarity.point_jd_interval, stationarity.py:39-56, the CPU check of a nominated
 * JD witness, witness_screen.py:79-88). out = NaN where the reference raises (a nonfinite or inverted step). */
int seqvp_point_jd_interval_host(const double* vertices, const double* spacing, const double* point, double* out);

/* HOST. Bernstein classification of one cell (jacobian_intervals + lagrange_to_bernstein + _classify with the
 * lowest-bound-first subdivision, feasibility.py:305-393). spacing = (hx, hy, hz). */
int seqvp_bernstein_cell(const double* vertices, const double* spacing, double delta, int max_depth, int64_t max_boxes,
                         SeqvpCellCertificate* out);

/* HOST. The fallback and final record of a screened map (feasibility.py:697-769): screen = the 5 summary values
 * of seqvp_certificate_screen (host copy), ncell = cells of the map, vertices (n, 8, 3) / cells (n, strictly
 * increasing) = the undecided cells. Jobs (one tile's cells each) run on `workers` threads. */
int seqvp_certificate_finish(const double* screen, int64_t ncell, const double* vertices, const int64_t* cells,
                             int64_t n, const double* spacing, double delta, int64_t tile, int max_depth,
                             int64_t max_boxes, int workers, SeqvpCertificate* out);

/* HOST. Edge-hull margins and flags of every cell of a host map u (nz, ny, nx, 3) on `workers` CPU threads, the
 * first stage of seqvp_certify_map_host (replaces parallel_margins, compiled_enclosure.py:193-227, 336-345):
 * margin[c] = lo(J_hull - delta), flag[c] = 1 when an interval step was nonfinite or inverted. Four cells of a row
 * per AVX2 evaluation when the CPU has AVX2; every unflagged margin is bitwise the scalar one. */
int seqvp_edge_hull_host(const double* u, int64_t nx, int64_t ny, int64_t nz, const double* spacing, double delta,
                         int workers, double* margin, uint8_t* flag);

/* HOST. The whole certificate of a host map u (nz, ny, nx, 3) on the CPU (certify_map, feasibility.py:395-431 and
 * _certify_map_parallel): nonfinite/boundary/domain checks against the axes ax/ay/az, edge hull, prescreen (D =
 * host derivative table, NULL = none), Bernstein fallback. The independent re-certification of an output. */
int seqvp_certify_map_host(const double* u, int64_t nx, int64_t ny, int64_t nz, const double* ax, const double* ay,
                           const double* az, const double* spacing, double delta, const double* D, int64_t tile,
                           int max_depth, int64_t max_boxes, int workers, SeqvpCertificate* out);

/* ---- FP32 pyramid pre-optimiser (replaces preoptimize, jd_curl/three_d/preopt.py:518-614, and the analytic
 *      gradient of preopt_grad.py) ----------------------------------------------------------------------------------
 * Fields: u (3, nz-2, ny-2, nx-2) component-major, voxel units, the interior nodes of a zero-face field; images
 * (nz, ny, nx). fp32 = 1: float32 buffers, 0: float64 (the error reference). */

/* Device workspace of seqvp_preopt_energy_grad. */
size_t seqvp_preopt_workspace_bytes(int64_t nx, int64_t ny, int64_t nz, int fp32);

/* P(u) and dP/du (_pre_energy_tensors, preopt.py:157-202; preopt_grad.py:120-205): trilinear border-clamped data
 * term, cell-centre (J - 1)^2 + |curl|^2 (weight mu/2), relu(j0 - J)^2 floor (weight kappa), seven-point curvature
 * (weight alpha/2). terms (6 device doubles) = [total, data, geometry, floor, curvature, min J]; grad like u.
 * Spacing (hx, hy, hz) of the level grid. */
int seqvp_preopt_energy_grad(const void* u, const void* fixed, const void* moving, int64_t nx, int64_t ny, int64_t nz,
                             double hx, double hy, double hz, double mu, double alpha, double kappa, double j0, int fp32,
                             void* grad, double* terms, void* workspace, size_t workspace_bytes, void* stream);

/* One update (the loop body of _run_level_default, preopt.py:219-233): grad smoothed by the separable kernel
 * w_grad[0..2 r_grad] (r_grad < 0: no smoothing; replicate padding), Adam moments m, v, step scaled so that
 * max |step| <= cap, u -= step, then u smoothed by w_field (r_field < 0: none). grad and tmp are overwritten.
 * state (2 device int64): [t, cap hits], both advanced on the device (t = iterations done). pmax: 1024 scratch
 * values of the field type. No host synchronisation: the launches can be captured in a graph. */
int seqvp_preopt_step(void* u, void* m, void* v, void* grad, void* tmp, int64_t nx, int64_t ny, int64_t nz,
                      const void* w_grad, int r_grad, const void* w_field, int r_field, double lr, double cap,
                      double beta1, double beta2, double eps, int fp32, long long* state, void* pmax, void* stream);

/* Pyramid resampling, float64 (downsample_array, nifti.py:128-149; prolong_displacement, multiresolution.py:55-84;
 * boundary_taper, preopt.py:617-624). Arrays are C-order (n0, n1, n2, nc).
 * seqvp_gauss_axis: correlation with w[0..2r] along axis 0/1/2, replicate padding (in != out).
 * seqvp_lerp_axis: out[.., j, ..] = (1 - weight[j]) in[.., left[j], ..] + weight[j] in[.., right[j], ..], m outputs
 * along the axis, separately rounded (bitwise the host prolongation). */
int seqvp_gauss_axis(const double* in, double* out, int64_t n0, int64_t n1, int64_t n2, int64_t nc, int axis,
                     const double* w, int r, void* stream);
int seqvp_lerp_axis(const double* in, double* out, int64_t n0, int64_t n1, int64_t n2, int64_t nc, int axis, int64_t m,
                    const int64_t* left, const int64_t* right, const double* weight, void* stream);
/* out = ((rz[z] ry[y]) rx[x]) u for a (nz, ny, nx, 3) field. */
int seqvp_taper(const double* u, int64_t nx, int64_t ny, int64_t nz, const double* rx, const double* ry,
                const double* rz, double* out, void* stream);
/* out = a u (n doubles). */
int seqvp_scale(const double* u, double a, double* out, int64_t n, void* stream);
/* Measurement probe with no counterpart in the reference (for the roofline of the FP64 kernels): `blocks`
 * workgroups of 256 lanes run `iters` rounds of 8 independent v_mfma_f64_16x16x4f64 (kind 0) or 8 independent FP64
 * vector FMAs per lane (kind 1) on registers only; out receives blocks * 256 doubles. The caller times the launch:
 * kind 0 performs blocks * 4 * iters * 8 * 2048 flop, kind 1 blocks * 256 * iters * 8 * 2. */
int seqvp_fp64_ceiling_probe(int kind, int64_t blocks, int64_t iters, double* out, void* stream);
/* (nz, ny, nx, 3) mm field <-> (3, nz-2, ny-2, nx-2) voxel-unit interior (float32 when fp32 = 1); the reverse
 * writes zero faces. */
int seqvp_field_to_interior(const double* u, int64_t nx, int64_t ny, int64_t nz, double hx, double hy, double hz,
                            int fp32, void* out, void* stream);
int seqvp_interior_to_field(const void* v, int64_t nx, int64_t ny, int64_t nz, double hx, double hy, double hz,
                            int fp32, double* out, void* stream);

/* HIP graphs of device-only solver segments (replaces the per-attempt host round trips of the reference's device
 * trial and direction, accelerated.py:1666-1730, hybrid.py:262-286, poisson_gradient.py:1241-1262, with one graph
 * launch). seqvp_graph_begin starts a thread-local capture on `stream` (not the null stream); the library calls
 * that follow on that stream are recorded, not run; seqvp_graph_end closes the capture and instantiates it into
 * *exec (SEQVP_ELAUNCH if a recorded call failed). seqvp_graph_launch runs it on any stream; seqvp_graph_destroy
 * frees it. A graph holds the device addresses it was captured with: the caller keys graphs by them. */
int seqvp_graph_begin(void* stream);
int seqvp_graph_end(void* stream, void** exec);
int seqvp_graph_launch(void* exec, void* stream);
int seqvp_graph_destroy(void* exec);
/* bytes from src to dst (device or page-locked host memory) on `stream`; recorded as a copy node under capture. */
int seqvp_copy_async(void* dst, const void* src, size_t bytes, void* stream);

#ifdef __cplusplus
}
#endif
#endif
