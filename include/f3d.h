/*
 * f3d.h -- C ABI of libf3d.so, the MI355X (gfx950) implementation of the
 * fiber3d GMRES-matvec hot path.
 *
 * Conventions (all entry points):
 *   - Every pointer argument is a DEVICE pointer owned by the caller, except
 *     where a comment says "host".
 *   - A 3-vector field over n points is addressed as p[i*sp + c*sc] for point
 *     i in [0,n) and component c in {0,1,2}:  AoS "xyzxyz" is (sp,sc)=(3,1),
 *     SoA "xx..yy..zz.." is (sp,sc)=(1,n).  All arrays of one side (sources,
 *     targets, output) share that side's (sp,sc).
 *   - Work is enqueued asynchronously on `stream` (a hipStream_t; NULL = the
 *     default stream).  No call synchronises the device.
 *   - Return 0 on success, a negative F3D_E* code on failure; the message is
 *     then available from f3d_last_error() (thread-local).
 *   - Scratch memory comes from the caller's `ws` buffer; the required size is
 *     given by the matching *_workspace_size() query.
 */
#ifndef F3D_H
#define F3D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------- */
#define F3D_OK 0
#define F3D_EINVAL (-1)    /* invalid argument (size, pointer, enum)      */
#define F3D_EHIP (-2)      /* a HIP runtime call failed                    */
#define F3D_EWORKSPACE (-3)/* workspace missing or too small               */

/* ---- pair-sum kernels ------------------------------------------------ */
#define F3D_OSEEN 0     /* Stokeslet: u_i = sum_j (1/(8 pi eta)) [f/r + d(d.f)/r^3] */
#define F3D_ROTLET 1    /* rotlet:    u_i = sum_j (1/(8 pi eta)) (tau x d)/r^3      */
#define F3D_STRESSLET 2 /* double layer: u_i = sum_j (-3/(4 pi eta)) (d.q)(d.n) d/r^5 */

#define F3D_F64 0  /* FP64 pair evaluation and accumulation                    */
#define F3D_F32C 1 /* FP32 pair evaluation, compensated (Kahan) per 64-source
                      tile, tile partials combined in FP64; output FP64        */

/* flags for f3d_pairsum */
#define F3D_ACCUMULATE 1u     /* out += result instead of out = result         */
#define F3D_CLAMP_NEAR 2u     /* F32C only: instead of the reference's
                                 regularisation r -> sqrt(r^2+reg^2) for
                                 r < eps, clamp 1/r at 1/eps (one instruction
                                 less per pair; identical results whenever no
                                 non-excluded pair is closer than eps)        */
#define F3D_EXCLUDE_SELF 4u   /* drop the pair (source i, target i): for a
                                 self-sum with the points as both sources
                                 and targets (the reference's j != i loop);
                                 not combinable with excl_lo/excl_hi          */

/* Returns a short description of the last error on the calling thread. */
const char* f3d_last_error(void);
/* ABI version (major*100 + minor). */
int f3d_abi_version(void);

/*
 * Launch shape of f3d_pairsum for one (kernel, precision): k targets per lane
 * (1, 2 or 4) and the number of wavefronts the source split aims for
 * (64 .. 2^20); 0 restores the default (2, 8192).  Process-wide, set by the
 * autotuner (fiber3d_amd/autotune.py); the workspace size depends on it, so
 * query f3d_pairsum_workspace_size after changing it.  No reference
 * counterpart (the numba kernels have no launch shape).
 */
int f3d_pairsum_set_launch(int kernel, int precision, int32_t k, int32_t target_waves);

/*
 * The same for one size bucket: calls with ceil(log2 nt) == lg_nt and
 * ceil(log2 ns) == lg_ns (both clamped to [8, 22]) use (k, target_waves);
 * k == 0 returns the bucket to the (kernel, precision) shape above.  The
 * autotuner fills every bucket from its nearest tuned problem size.  No
 * reference counterpart.
 */
int f3d_pairsum_set_launch_bucket(int kernel, int precision, int32_t lg_nt, int32_t lg_ns,
                                  int32_t k, int32_t target_waves);

/*
 * The launch plan f3d_pairsum would use for (kernel, precision, ns, nt):
 * out[0] = targets per lane, out[1] = source chunks, out[2] = sources per
 * chunk.  Host-only (no device work); no reference counterpart.
 */
int f3d_pairsum_plan(int kernel, int precision, int64_t ns, int64_t nt, int32_t* out);

/*
 * Launch choices of the batched LU for one precision (F3D_F64 or F3D_F32C):
 * waves per CU of the one-wave blocked factorisation (1..4 for FP64, 1..8
 * for FP32; 0 restores the largest), and whether solves without refinement
 * stage the factor in LDS when it fits in 64 KB (1, the default) or stream
 * it from HBM (0).  Set by the autotuner; no reference counterpart.
 */
int f3d_lu_set_launch(int precision, int32_t factor_waves_per_cu, int32_t solve_staged);

/*
 * The kernel f3d_gemm_bf16 uses for N == n (128 or 256, row-major operands,
 * K <= 256): enable 2 (the default) the resident-weight kernel (B held in
 * LDS by one persistent workgroup per CU; K a multiple of 16, A rows 16-byte
 * aligned, else the streaming kernel), 1 the streaming kernel, 0 the tiled
 * kernel.  Set by the autotuner; no reference counterpart.
 */
int f3d_gemm_set_stream(int32_t n, int32_t enable);

/*
 * Timing of the pair kernel alone (not the bounding-box, pack and reduce
 * launches around it): enable != 0 clears the record and makes every later
 * f3d_pairsum / f3d_periphery_apply record a HIP event pair around its pair
 * kernel on the caller's stream (at most 4096 launches are kept); enable == 0
 * stops recording.  f3d_pairsum_timer_read waits for the last recorded launch
 * and returns the summed milliseconds and the launch count.  Process-wide,
 * one device, not thread safe.  Instrumentation for benchmarks and profilers;
 * no reference counterpart.
 */
int f3d_pairsum_timer(int enable);
int f3d_pairsum_timer_read(double* total_ms, int32_t* launches);

/* Bytes of workspace f3d_pairsum needs for these sizes. */
size_t f3d_pairsum_workspace_size(int kernel, int precision, int64_t ns, int64_t nt);

/*
 * Source-target pair sum, matrix free.  Replaces
 *   kernels.oseen_kernel_source_target_numba      (python/kernels/kernels.py:794-846)
 *   kernels.rotlet_kernel_source_target_numba     (python/kernels/kernels.py:856-908)
 *   kernels.stresslet_kernel_source_target_numba  (python/kernels/kernels.py:1186-1217)
 *   kernels.stresslet_kernel_times_normal_times_density_numba
 *                                                 (python/kernels/kernels.py:1123-1183;
 *      call with the same points as sources and targets and F3D_EXCLUDE_SELF,
 *      the reference's j != i loop)
 * and, with src_w != NULL, the weight multiply of tstep_utils.flow_fibers
 *   (python/tstep/tstep_utils.py:653-700).
 *
 *   d = t_i - s_j, r = |d|; if r < eps then r <- sqrt(r^2 + reg^2) and d is
 *   kept (numba semantics, kernels.py:830-831).
 *
 *   src_pos   [ns] positions of the sources
 *   src_nrm   [ns] normals (STRESSLET only, else NULL)
 *   src_den   [ns] densities f / tau / q
 *   src_w     [ns] optional scalar weight per source (density is multiplied)
 *   trg_pos   [nt] target positions
 *   excl_lo/excl_hi [nt] optional int32 half-open source-index range
 *             [lo_i, hi_i) whose pairs are EXCLUDED for target i (same-fiber
 *             exclusion, or lo=i,hi=i+1 to drop the self term); both NULL or
 *             both set.  Replaces the add-then-subtract of the per-fiber self
 *             block (tstep_utils.py:703-729, :812-851).
 *   out       [nt] result, scaled by the kernel constant
 */
int f3d_pairsum(int kernel, int precision, uint32_t flags,
                int64_t ns, const double* src_pos, const double* src_nrm,
                const double* src_den, const double* src_w, int64_t s_sp, int64_t s_sc,
                int64_t nt, const double* trg_pos, int64_t t_sp, int64_t t_sc,
                const int32_t* excl_lo, const int32_t* excl_hi,
                double eta, double reg, double eps,
                double* out, int64_t o_sp, int64_t o_sc,
                void* ws, size_t ws_bytes, void* stream);

/*
 * Allocation-free f3d_pairsum.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results as f3d_pairsum, without ws/ws_bytes: the scratch is a static
 * device array of F3D_NOALLOC_SCRATCH_BYTES per GPU, held from the library's
 * first call on that GPU until the process ends.  Where the default launch
 * shape needs more, the sources are split into fewer chunks (the summation
 * order, hence the rounding, may then differ from f3d_pairsum's); sizes that
 * do not fit even one chunk (source records plus 24 bytes per target) are
 * rejected with F3D_EWORKSPACE.  Nothing is allocated or freed.  Every call
 * on one GPU shares that GPU's scratch: calls that overlap on one GPU (two
 * streams, two threads) race and must be serialised by the caller.
 */
#define F3D_NOALLOC_SCRATCH_BYTES (64u << 20)
int f3d_pairsum_noalloc(int kernel, int precision, uint32_t flags,
                        int64_t ns, const double* src_pos, const double* src_nrm,
                        const double* src_den, const double* src_w, int64_t s_sp, int64_t s_sc,
                        int64_t nt, const double* trg_pos, int64_t t_sp, int64_t t_sc,
                        const int32_t* excl_lo, const int32_t* excl_hi,
                        double eta, double reg, double eps,
                        double* out, int64_t o_sp, int64_t o_sc, void* stream);

/* ---- periphery (confining shell) operator ---------------------------- */

size_t f3d_periphery_workspace_size(int precision, int64_t n);

/*
 * Matrix-free periphery double-layer operator, replacing the dense
 * shell_stresslet + shell_complementary matvecs of
 *   tstep.A_body_fiber_hyd (python/tstep/tstep.py:1097-1104), whose matrices
 *   are built at python/tstep/tstep.py:328-345 and
 *   periphery.Periphery.get_singularity_subtraction_vectors
 *   (python/periphery/periphery.py:75-98):
 *
 *   out_i = sum_{j != i} S(x_i - x_j; n_j) q_j
 *           - sum_c (q_{i,c} / w_i) e_{c,i} - q_i / w_i + n_i sum_j n_j . q_j
 *           [+ add_i]
 *
 *   pos, nrm, q, ex, ey, ez, add, out: [n] 3-vectors with the same (sp,sc);
 *   w: [n] quadrature weights; add may be NULL.
 */
int f3d_periphery_apply(int precision, int64_t n, const double* pos, const double* nrm,
                        const double* w, const double* ex, const double* ey,
                        const double* ez, const double* q, const double* add,
                        int64_t sp, int64_t sc, double eta, double reg, double eps,
                        double* out, void* ws, size_t ws_bytes, void* stream);

/*
 * Dense Oseen tensor, row-major with leading dimension ld: block (i, j) =
 * (I/r + d d^T/r^3)/(8 pi eta) between target i and source j, d = t_i - s_j,
 * r regularised to sqrt(r^2 + reg^2) where r <= eps.  Point k of the sources
 * is (src[k s_sp], src[k s_sp + s_sc], src[k s_sp + 2 s_sc]) (AoS: 3, 1;
 * 'xyz' blocks: 1, ns), likewise the targets.  layout 0: entry (3i+a, 3j+b)
 * ('r' format); layout 1: entry (a nt + i, b ns + j) ('xyz' format).
 * Replaces kernels.oseen_tensor and kernels.oseen_tensor_source_target
 * (python/kernels/kernels.py:31-116, :119-211), used for the per-fiber self
 * blocks (python/tstep/tstep_utils.py:717-722, :794-800).
 */
int f3d_oseen_matrix(int64_t ns, const double* src, int64_t s_sp, int64_t s_sc, int64_t nt,
                     const double* trg, int64_t t_sp, int64_t t_sc, double eta, double reg,
                     double eps, int32_t layout, double* M, int64_t ld, void* stream);

/*
 * Dense 3n x 3n stresslet-times-normal matrix, row-major with leading
 * dimension ld: block (i, j != i) = (-3/(4 pi eta)) (d . n_j)/r^5 d d^T with
 * d = x_i - x_j (regularised for r < eps), zero diagonal blocks.  Replaces
 * kernels.stresslet_kernel_times_normal_numba (python/kernels/kernels.py:1063-1120),
 * used to build the dense shell preconditioner (python/tstep/tstep.py:329).
 */
int f3d_stresslet_matrix(int64_t n, const double* pos, const double* nrm, int64_t sp,
                         int64_t sc, double eta, double reg, double eps, double* M, int64_t ld,
                         void* stream);

/* ---- batched dense LU over blocks of uneven size --------------------- */
/*
 * A batch of `nb` square blocks; block b has order n[b] <= 256 and is stored
 * column-major at A + a_off[b] (leading dimension n[b]).  n, a_off, x_off and
 * order are DEVICE arrays (int32, int64, int64, int32).  `order` (nullable)
 * lists the blocks in the sequence the persistent workgroups pick them up
 * (largest first balances uneven sizes); NULL means 0..nb-1.
 *
 * f3d_lu_factor_batched replaces tstep_utils.build_block_diagonal_lu_
 * preconditioner_fiber (python/tstep/tstep_utils.py:1405-1419), i.e.
 * scipy.linalg.lu_factor (LAPACK dgetrf, partial pivoting, first maximal
 * |pivot| wins ties) per fiber, in ONE persistent launch.
 *   precision F3D_F64: A is factored in FP64 into lu64 (offsets a_off).
 *   precision F3D_F32C: A is first equilibrated by powers of two,
 *     As = diag(rs) A diag(cs), then factored in FP32 into lu32 (offsets
 *     a_off); rs, cs are written at offsets x_off.  Solves refine in FP64.
 *   piv  [x_off]: LAPACK-style 0-based row interchanges (row k <-> piv[k]).
 *   perm [x_off]: the resulting row permutation, row k of LU is row perm[k]
 *                 of A (used by the solves).
 *   info [nb]: 0, or k+1 for the first exactly-zero pivot U(k,k).
 */
int f3d_lu_factor_batched(int precision, int64_t nb, int32_t nmax, const int32_t* n,
                          const int64_t* a_off, const int64_t* x_off, const int32_t* order,
                          const double* A, double* lu64, float* lu32, double* rs, double* cs,
                          int32_t* piv, int32_t* perm, int32_t* info, void* stream);

/*
 * Solve A_b y_b = x_b for every block (x, y at offsets x_off; y may alias x).
 * Replaces the per-fiber scipy.linalg.lu_solve loop of tstep.P_inv
 * (python/tstep/tstep.py:985-988) with one launch, one wavefront per block.
 *   F3D_F64: one forward/backward substitution with lu64.
 *   F3D_F32C: substitution with the FP32 factors (arithmetic in FP64), then up
 *     to `refine` sweeps of FP64 iterative refinement, r = x - A y with the
 *     FP64 matrix A, stopping per block once
 *     ||r||_inf <= tol * ||A||_inf * ||y||_inf.
 *   iters [nb] (nullable): refinement sweeps used per block.
 */
int f3d_lu_solve_batched(int precision, int64_t nb, int32_t nmax, const int32_t* n,
                         const int64_t* a_off, const int64_t* x_off, const double* A,
                         const double* lu64, const float* lu32, const double* rs,
                         const double* cs, const int32_t* perm, const double* x, double* y,
                         int32_t refine, double tol, int32_t* iters, void* stream);

/*
 * Per-fiber assembly of the implicit fiber system for nfib fibers of one node
 * count N (5 <= N <= 64), FP64, one workgroup per fiber.  Fiber f starts at
 * node node0[f] of the AoS positions X (npts x 3) and has length length[f].
 * D0: the four derivative matrices of the unit parameter interval, (4, N, N)
 * row-major; PX ((N-4) x N) and PT ((N-2) x N) the projections onto the
 * interior collocation points; w0 (N) the unit quadrature weights.  Writes:
 *   A_out + a_off[f]: A_sq (4N x 4N, column-major): the projected linear
 *     operator plus the 14 force- and torque-free end conditions;
 *   F_out + f_off[f]: the three force-operator row blocks (N x 4N each,
 *     column-major, adjacent);
 *   Q0/Q1/Q2 + q_off[3f + c]: the tension projections ((4N-14) x N);
 *   xs_out, xss_out (npts x 3) and w_out (npts) at the fiber's nodes.
 * Replaces fiber.form_linear_operator (python/fiber/fiber.py:322-447),
 * apply_BC_rectangular (:810-961), force_operator (:1399-1431) and the
 * per-fiber tension projection of python/tstep/tstep_utils.py:1238-1336.
 */
int f3d_fiber_assemble(int64_t nfib, int32_t N, const int64_t* node0, const double* length,
                       const int64_t* a_off, const int64_t* f_off, const int64_t* q_off,
                       const double* X, const double* D0, const double* PX, const double* PT,
                       const double* w0, double E, double dt, double c0, double c1,
                       double penalty, double* A_out, double* F_out, double* Q0, double* Q1,
                       double* Q2, double* xs_out, double* xss_out, double* w_out, void* stream);

/*
 * Per-fiber right-hand side rows for nfib fibers of one node count N
 * (5 <= N <= 64), FP64, one workgroup per fiber, after f3d_fiber_assemble
 * (xs, xss at the nodes) and the explicit flow v (npts x 3, AoS) of the
 * plus-end forces F_end x_s(L).  Writes the fiber's 4N rows at
 * rhs + rhs_off[f]: the projected collocation rows P R and the 14 end rows
 * (the plus-end force and its tangential part).  D0 as in
 * f3d_fiber_assemble (only D_1 is read).  Replaces fiber.compute_RHS with the
 * plus-end force (python/fiber/fiber.py:473-582), the RHS part of
 * apply_BC_rectangular (:810-961) and the per-fiber RHS assembly of
 * python/tstep/tstep.py:794-846 and tstep_utils.py:1238-1336.
 */
int f3d_fiber_rhs(int64_t nfib, int32_t N, const int64_t* node0, const double* length,
                  const int64_t* rhs_off, const double* X, const double* xs, const double* xss,
                  const double* v, const double* D0, const double* PX, const double* PT,
                  double dt, double c0, double c1, double penalty, double end_force,
                  double* rhs, void* stream);

/*
 * y_b = alpha * M_b x_b + beta * y_b for a batch of dense COLUMN-major blocks
 * of shape m[b] x k[b] (leading dimension m[b]) stored at M + m_off[b]; x at
 * x_off[b], y at y_off[b] (y must not alias x).  m[b] <= mmax <= 1024.
 * Replaces the block-diagonal CSR apply As_fibers.dot(XT)
 * (python/tstep/tstep.py:1173; python/multi_fibers/multi_fiber.py:47-78) and
 * the per-fiber force operator / tension projection products
 * (python/tstep/tstep.py:1036, :1160, :1171).
 */
int f3d_block_gemv_batched(int64_t nb, int32_t mmax, const int32_t* m, const int32_t* k,
                           const int64_t* m_off, const int64_t* x_off, const int64_t* y_off,
                           const double* M, const double* x, double* y, double alpha,
                           double beta, void* stream);

/*
 * y_b = M_b x_b for `count` dense ROW-major m x m FP64 matrices stored
 * back to back (M + b*m*m), x and y of length count*m (y must not alias x).
 * Applies the shell preconditioner's dense inverse, the same linear map as the
 * reference's lu_solve with the shell LU (python/tstep/tstep.py:346-351,
 * :957-958), and the rigid bodies' precomputed inverses
 * (python/tstep/tstep.py:225-252, :966-980).
 */
int f3d_dense_gemv_batched(int64_t count, int32_t m, const double* M, const double* x,
                           double* y, void* stream);

/*
 * The same with FP32 matrix entries (FP64 x, products, sums and y): half the
 * HBM traffic of the FP64 sweep.  Applies the shell inverse as the fast
 * mode's preconditioner, beside the FP32 fiber LU factors; the reference's
 * shell lu_solve (python/tstep/tstep.py:957-958) is the map it approximates.
 */
int f3d_dense_gemv_batched_f32(int64_t count, int32_t m, const float* M, const double* x,
                               double* y, void* stream);

/* ---- restarted GMRES on the device ----------------------------------- */
/*
 * Replace the host side of the reference's GMRES (SciPy's restarted gmres
 * under bpm_utilities/gmres.py:94-108, called at python/tstep/tstep.py:1209-
 * 1216): the Arnoldi orthogonalisation (classical Gram-Schmidt with one
 * re-orthogonalisation), the Hessenberg matrix, its Givens rotations and the
 * least-squares residual estimate, all FP64 in device memory.  n is the
 * system size, m (1..159) the restart length; V is the (m+1) x n basis
 * (row k = basis vector k, contiguous); ws is zero-filled once, of
 * f3d_gmres_workspace_size(n, m) bytes, and holds the state of one solve at
 * a time, including the current step j, so a sequence of calls can be
 * recorded in a HIP graph and replayed for every step.
 *   begin:   beta = ||r||, V_0 = r / beta, g = beta e_1, j = 0.
 *   current: xin = V_j (the operator's input).
 *   arnoldi: given w = A P^-1 V_j, orthogonalise w against V_0..V_j (CGS2),
 *            store column j of the Hessenberg matrix (rotated and not),
 *            rotate, update g, record est_j = |g_{j+1}| (ws, see gmres.hip),
 *            V_{j+1} = w / h_{j+1,j}, j += 1.  w is overwritten.
 *   solve:   y = H_k^-1 g_k over the first k steps, upd = V_k y; with r
 *            non-NULL also r_upd = r - V_{k+1} (H_raw y), the residual the
 *            Arnoldi relation gives for the correction.
 */
size_t f3d_gmres_workspace_size(int64_t n, int32_t m);
int f3d_gmres_begin(int64_t n, int32_t m, const double* r, double* V, void* ws, void* stream);
int f3d_gmres_current(int64_t n, int32_t m, const double* V, double* xin, void* ws,
                      void* stream);
int f3d_gmres_arnoldi(int64_t n, int32_t m, double* V, double* w, void* ws, void* stream);
int f3d_gmres_solve(int64_t n, int32_t m, int32_t k, const double* V, double* upd,
                    const double* r, double* r_upd, void* ws, void* stream);

/* ---- BF16 MFMA GEMM for the learned initial-guess network ------------ */
/*
 * No counterpart in the reference (the learned GMRES initial guess is new).
 * C = epi(alpha * A B) with A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn]
 * (FP32 in memory, rounded to BF16 for v_mfma_f32_32x32x16_bf16, FP32
 * accumulation), C(m,n) = C[m*scm + n*scn].  epi:
 *   0 store; 1 + bias[n], store pre-activation to Z (nullable), ReLU;
 *   2 multiply by ReLU'(Zin) (Zin > 0); 3 atomic add (C pre-zeroed; the only
 *   mode allowed with ksplit > 1); 4 + bias[n].
 */
int f3d_gemm_bf16(int epi, int64_t M, int64_t N, int64_t K, int32_t ksplit, const float* A,
                  int64_t sam, int64_t sak, const float* B, int64_t sbk, int64_t sbn, float* C,
                  int64_t scm, int64_t scn, const float* bias, float* Z, const float* Zin,
                  float alpha, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* F3D_H */
