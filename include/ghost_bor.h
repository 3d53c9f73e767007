/*
 * libghost_bor: MI355X (gfx950) body-of-revolution MoM hot path.
 *
 * C ABI only (plain pointers, sizes and ints; no torch types).  Every entry
 * point returns 0 on success or a GHOST_ERR_* code, with the message
 * available from ghost_last_error() (thread-local), except the three
 * void-returning compat samplers whose signatures are fixed by the reference.
 *
 * Complex arrays are interleaved (re, im) doubles, row-major.
 */
#ifndef GHOST_BOR_H
#define GHOST_BOR_H

#ifdef __cplusplus
extern "C" {
#endif

#define GHOST_FP64 0  /* FP64 fill + FP64 LU */
#define GHOST_MIXED 1 /* FP32 far fill + FP32 LU, FP64 near band, residual, refinement, far field */

#define GHOST_UNMASKED 1 /* add_problem flag: keep all 2 Nn bases (no axis / open-end mask) */

#define GHOST_ERR_ARG 1
#define GHOST_ERR_STATE 2
#define GHOST_ERR_HIP 3
#define GHOST_ERR_OOM 4

typedef struct ghost_bor ghost_bor;

/* Message of the last failing call on this thread. */
const char* ghost_last_error(void);

/* ── compat samplers ─────────────────────────────────────────────────────
 * Replace the ctypes-bound natives of bor_stream_kernel.c (loaded by
 * bor_streaming.py:50-76, called at bor_streaming.py:289-293, 328-331):
 * identical names, argument order and meaning; host pointers in/out,
 * synchronous, reentrant; outputs are caller-allocated complex128
 * [nr, np_, nxi] viewed as interleaved doubles; R is clamped at 1e-30.
 * The void forms keep the reference ABI (no error channel); the ghost_
 * forms return an error code. */
void sample_g(int nr, int np_, int nxi, const double* rho_p, const double* z_p, const double* rho_q,
              const double* z_q, double k, const double* sin2_tab,
              double* out); /* bor_stream_kernel.c:26-48 */
void sample_mfie(int nr, int np_, int nxi, const double* rho_p, const double* z_p, const double* tr_p,
                 const double* tz_p, const double* rho_q, const double* z_q, const double* tr_q,
                 const double* tz_q, double k, const double* cx_tab, const double* sx_tab, double* o_tt,
                 double* o_tf, double* o_ft, double* o_ff); /* bor_stream_kernel.c:65-106 */
void sample_ibc(int nr, int np_, int nxi, const double* rho_p, const double* z_p, const double* tr_p,
                const double* tz_p, const double* rho_q, const double* z_q, const double* tr_q,
                const double* tz_q, double k, const double* cx_tab, const double* sx_tab, double* o_tt,
                double* o_tf, double* o_ft, double* o_ff); /* bor_stream_kernel.c:108-150 */
int ghost_sample_g(int nr, int np_, int nxi, const double* rho_p, const double* z_p, const double* rho_q,
                   const double* z_q, double k, const double* sin2_tab, double* out);
int ghost_sample_mfie(int nr, int np_, int nxi, const double* rho_p, const double* z_p, const double* tr_p,
                      const double* tz_p, const double* rho_q, const double* z_q, const double* tr_q,
                      const double* tz_q, double k, const double* cx_tab, const double* sx_tab,
                      double* o_tt, double* o_tf, double* o_ft, double* o_ff);
int ghost_sample_ibc(int nr, int np_, int nxi, const double* rho_p, const double* z_p, const double* tr_p,
                     const double* tz_p, const double* rho_q, const double* z_q, const double* tr_q,
                     const double* tz_q, double k, const double* cx_tab, const double* sx_tab,
                     double* o_tt, double* o_tf, double* o_ft, double* o_ff);

/* ghost_sample_mfie with no device allocation: the same arguments, meaning
 * and results (bitwise), but inputs and outputs are staged through a
 * file-scope device array of GHOST_NOALLOC_SCRATCH_BYTES per device instead of
 * buffers allocated per call, and the work runs on the device's null stream.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (tight loops, harnesses that forbid hipMalloc).  Limit: the staged
 * bytes, 8 (4 nr + 4 np_ + 2 nxi) + 64 nr np_ nxi with each of the 14 pieces
 * rounded up to 256, must not exceed GHOST_NOALLOC_SCRATCH_BYTES (else
 * GHOST_ERR_ARG; split the rows).  Calls that overlap on one device share
 * the array and race. */
#define GHOST_NOALLOC_SCRATCH_BYTES (48u << 20)
int ghost_sample_mfie_noalloc(int nr, int np_, int nxi, const double* rho_p, const double* z_p,
                              const double* tr_p, const double* tz_p, const double* rho_q, const double* z_q,
                              const double* tr_q, const double* tz_q, double k, const double* cx_tab,
                              const double* sx_tab, double* o_tt, double* o_tf, double* o_ft, double* o_ff);

/* ── device-resident mode sweep ──────────────────────────────────────────
 * One handle holds a batch of problems (generatrix x frequency), each with
 * its 2 m_max + 1 mode systems, on the device current at creation and on a
 * stream of its own.  Replaces the per-frequency body of solve_bor
 * (bor_solver.py:820-1004): BorPecSolver's fill (293-345, 399-435), the
 * excitation (625-645, 437-461), _mode_sweep's np.linalg.solve (744-758)
 * and farfield_mode (659-708).  Calls on one handle must not overlap. */

/* precision: GHOST_FP64 or GHOST_MIXED; nb: LU block size (multiple of 8, <= 64; 0 = 64). */
int ghost_bor_create(ghost_bor** out, int precision, int nb);
int ghost_bor_destroy(ghost_bor* h);

/* Keep FP64 copies of A and B for ghost_bor_residual (default on; before build). */
int ghost_bor_set_residual(ghost_bor* h, int on);

/* Mode pairing (default on; before build).  The reference solves every mode
 * system on its own (bor_solver.py:744-753).  Here Z_{-m} = S Z_m S, with S
 * the identity on the J_t rows and minus it on the J_phi rows, and the
 * plane-wave right-hand sides obey S b_{-m} = [b_m(VV), -b_m(HH)], so
 * x_{-m} = S [x_m(VV), -x_m(HH)]: only modes m >= 0 are filled, factored and
 * solved, and mode -m reads mode m's storage with those signs.  Accessors by
 * (problem, mode) are unchanged; ghost_bor_set_solution of mode -m sets mode
 * m to the mirror image.  No counterpart in the reference; ghost_bor_adjoint
 * needs a handle built with pairing off. */
int ghost_bor_set_pairing(ghost_bor* h, int on);

/* Add one problem.  nodes: [Nn][2] (rho, z), traversed +z end first
 * (bor_kernels.py:38-81).  k: real wavenumber (2 pi f / c0).  m_max: modes
 * -m_max..m_max.  n_xi_efie / n_xi_bracket: azimuth FFT sizes (powers of two;
 * bor_kernels.py:149-152, 315-319).  Z = c_efie Z_EFIE + c_mfie Z_MFIE and
 * V = c_efie V_E + c_mfie V_H (bor_solver.py:929-946: CFIE c_efie = alpha,
 * c_mfie = (1 - alpha) eta0).  thetas_deg: monostatic aspects from +z.
 * flags: 0 or GHOST_UNMASKED.  *idx receives the problem index. */
int ghost_bor_add_problem(ghost_bor* h, const double* nodes, int Nn, double k, int m_max, int n_xi_efie,
                          int n_xi_bracket, double c_efie, double c_mfie, const double* thetas_deg,
                          int n_theta, int flags, int* idx);

/* Surface impedance of problem prob (before build): zs complex [Ne]
 * (interleaved), one Leontovich Z_s (ohm) per element.  Adds the IBC-EFIE
 * operator c_mass (1/2) <W, Z_s J> + c_ibc <W, PV curl int G M> with the
 * eliminated magnetic current M = -Z_s n x J (bor_solver.py:463-552;
 * solve_bor's assemble closure, bor_solver.py:929-934, uses c_ibc = c_mass =
 * alpha), and the far field radiates M as well (bor_solver.py:697-707).
 * c_mass = 0 with zs = 1 leaves the bare rotated principal-value operator
 * (the PMCHWT blocks, bor_solver.py:554-583).  The impedance operator takes
 * the MFIE brackets' place in the fill: c_mfie must be 0 when c_ibc != 0 (the
 * reference refuses CFIE with Z_s, bor_solver.py:867-869).  All-zero zs
 * leaves the problem a PEC one. */
int ghost_bor_set_impedance(ghost_bor* h, int prob, const double* zs, double c_ibc, double c_mass);

/* Homogeneous medium of problem prob's fill (before build): wavenumber
 * k + j k_im (k from ghost_bor_add_problem, k_im <= 0 for a lossy medium)
 * and wave impedance eta_re + j eta_im (ohm); the EFIE operator is then
 * j k eta (A - grad Phi / k^2) with the medium's Green function
 * (bor_solver.py:175-199, medium=(eps, mu); bor_kernels.py:96-99 complex k).
 * k_im != 0 needs a GHOST_FP64 handle.  Free space by default. */
int ghost_bor_set_medium(ghost_bor* h, int prob, double k_im, double eta_re, double eta_im);

/* Make problem prob a closed homogeneous penetrable body (PMCHWT,
 * bor_solver.py:1019-1094, solve_bor_dielectric): the problem, added with
 * c_efie = 1 and c_mfie = 0 in free space, is the exterior; the interior
 * medium has wavenumber k_re + j k_im (k_im <= 0) and wave impedance
 * eta_re + j eta_im, filled on its own azimuth grids n_xi_efie /
 * n_xi_bracket.  Each mode system becomes
 *   [[Te + Ti, -eta0 (Pe + Pi)], [eta0 (Pe + Pi), Te + (eta0/eta_i)^2 Ti]]
 * in (J, M / eta0), masked per half (twice the single-surface size), its
 * right-hand sides [<W, E_inc>; eta0 <W, H_inc>], and the far field
 * radiates J and M.  GHOST_FP64 handles only; before build. */
int ghost_bor_set_interior(ghost_bor* h, int prob, double k_re, double k_im, double eta_re, double eta_im,
                           int n_xi_efie, int n_xi_bracket);

/* ── Multi-region bodies (PEC cores under coating layers, coating patches;
 * bor_solver.py:1100-1297, 1805-2172).  FP64 handles only.  A body is
 * described by operators, cross operators and one region set:
 *
 * ghost_bor_add_operator: one generatrix (nodes [Nn][2], +z end first) in
 * one homogeneous medium (k = k_re + j k_im, k_im <= 0; eta), filled as
 * operators only: its EFIE operator T and, brackets != 0, the unit-weight
 * rotated-PV brackets of its coupling operator P (bor_solver.py:554-575,
 * BorPecSolver.assemble_mode / assemble_pmchwt_P of that medium).  thetas:
 * the aspects, for a free-space surface that bounds the exterior (n_theta may
 * be 0 otherwise). */
int ghost_bor_add_operator(ghost_bor* h, const double* nodes, int Nn, double k_re, double k_im, double eta_re,
                           double eta_im, int m_max, int n_xi_efie, int n_xi_bracket, int brackets,
                           const double* thetas_deg, int n_theta, int* op);
/* The cross-surface T and P between the test bases of op_test and the
 * source bases of op_src (same medium, same m_max; BorCrossOperators,
 * bor_solver.py:1115-1297): element pairs sharing an endpoint on corner
 * cells, pairs closer than near_factor x the longer element on a
 * near_order^2 tensor grid (the reference's defaults 2.0 and 12). */
int ghost_bor_add_cross(ghost_bor* h, int op_test, int op_src, double near_factor, int near_order, int* cross);
/* A region set: per mode -m_max..m_max a system over n_full unknowns before
 * the junction map, at aspects thetas_deg. */
int ghost_bor_add_region(ghost_bor* h, int m_max, int n_full, const double* thetas_deg, int n_theta, int* set);
/* Add (c_re + j c_im) x block at (row_off, col_off) of every mode's full
 * system: src 0 operator idx, 1 cross operator idx; which 0 T, 1 P
 * (_MultiRegionBor.assemble, bor_solver.py:1988-2015). */
int ghost_bor_region_block(ghost_bor* h, int set, int src, int idx, int which, double c_re, double c_im,
                           int row_off, int col_off);
/* Junction map of mode m (_MultiRegionBor.build_Q, bor_solver.py:1908-1986):
 * full index i -> reduced index red_of_full[i] (-1: dropped) with real
 * coefficient coef[i]; the mode's system is Q^T A Q (n_red <= 8192). */
int ghost_bor_region_map(ghost_bor* h, int set, int m, int n_red, const int* red_of_full, const double* coef);
/* A free-space operator surface bounding the exterior: its J rows start at
 * off_J of the full system and, off_M >= 0, its M / eta0 rows follow at
 * off_M = off_J + 2 Nn.  Its rows take <W, E_inc> (and eta0 <W, H_inc>) and
 * its currents radiate (bor_solver.py:2017-2037). */
int ghost_bor_region_exterior(ghost_bor* h, int set, int op, int off_J, int off_M);
/* Far field of a region set, complex [2m_max+1][2][n_theta] (as get_dF). */
int ghost_bor_region_get_dF(ghost_bor* h, int set, double* out);
/* One mode of a region set: *n its size; which 0: A [n][n] and B [n][nrhs]
 * after fill; which 1: X [n][nrhs] after solve (complex interleaved). */
int ghost_bor_region_system(ghost_bor* h, int set, int m, int which, int* n, double* out1, double* out2);

/* Geometry tables, device buffers and the launch schedule. */
int ghost_bor_build(ghost_bor* h);

/* Fill: near band, far blocks, masked mode systems and RHS columns. */
int ghost_bor_fill(ghost_bor* h);
/* Batched LU + solve of every mode system.  GHOST_MIXED: refine_iters >= 0
 * runs that many FP64-residual refinement sweeps; refine_iters < 0 is
 * adaptive: before each of up to -refine_iters sweeps an error probe
 * ||A32^-1 (B s - A X s)|| / ||X s|| (two fixed random column combinations
 * s; residual in FP64, correction with the FP32 factors), the estimated
 * relative error of each system's currents, is measured and the sweeps stop
 * once its maximum over systems is <= the probe tolerance.  The FP32
 * right-hand sides carry power-of-two column scales, so orders whose
 * excitation lies below the FP32 range are solved too. */
int ghost_bor_solve(ghost_bor* h, int refine_iters);
/* Tile variant of the batched complex MFMA GEMM (trailing updates, solves,
 * residuals): 0 = 64x64 single-buffered, 1 = 64x64, 2 = 64x128, 3 = 128x128
 * (FP32; 64x128 for FP64) with double-buffered K chunks, 4 = variant 1 with
 * the C tile loaded before the K loop; 5 (FP32 factors; FP64 handles run 4)
 * = 64x64 on the bf16 MFMA with every FP32 operand split exactly into three
 * bf16 parts (rounded to nearest) and the six leading cross products
 * accumulated in FP32 (FP32 accuracy); -1 (default) = 5 for FP32, 4 for FP64. */
int ghost_bor_set_gemm_variant(ghost_bor* h, int variant);
/* Panels per LU super-block (before build; default 8 for FP32 factors, 4 for FP64): the trailing update
 * runs once per super-block with K = nb x panels, the panels and row blocks
 * inside it catch up through narrow GEMMs.  1 = plain right-looking LU. */
int ghost_bor_set_lu_super(ghost_bor* h, int panels);
/* LU and back substitution on 2 streams (default: two interleaved halves of
 * the systems, one half's panel phases beside the other's trailing updates)
 * or 1 (the whole batch in one schedule).  No counterpart in the reference. */
int ghost_bor_set_lu_streams(ghost_bor* h, int n);
/* Probe tolerance of the adaptive refinement (default 1e-5). */
int ghost_bor_set_probe_tol(ghost_bor* h, double tol);
/* Probe errors of the last adaptive solve, one per probe; *n: capacity in, count out. */
int ghost_bor_refine_history(ghost_bor* h, double* out, int* n);
/* Monostatic far-field contribution of every mode and column. */
int ghost_bor_farfield(ghost_bor* h);
/* fill + solve + farfield. */
int ghost_bor_run(ghost_bor* h, int refine_iters);
int ghost_bor_sync(ghost_bor* h);
/* "name|gfx arch|cuN" of the current device (NUL-terminated, truncated to
 * len): the key under which tuned tile choices are cached.  No counterpart
 * in the reference (part of the tuner query). */
int ghost_device_name(char* buf, int len);
/* Device-wide barrier (hipDeviceSynchronize) on the current device. */
int ghost_device_synchronize(void);

/* Capture fill + solve + far field of the whole batch into one hipGraph, and
 * replay it (asynchronously) on the handle's stream; graph_ms gives the last
 * replay's time.  Together they replace the reference's CPU mode-sweep
 * driver loop (bor_solver.py:774-796) for a whole batch of problems. */
int ghost_bor_capture(ghost_bor* h, int refine_iters);
int ghost_bor_launch(ghost_bor* h);
int ghost_bor_graph_ms(ghost_bor* h, float* ms);

/* dF of one problem: complex [2 m_max + 1][2 (VV, HH)][n_theta], mode index m + m_max. */
int ghost_bor_get_dF(ghost_bor* h, int prob, double* out);
/* Number of systems of the handle (every problem's modes and the dense
 * systems, in the order added): the length of the per-system arrays below. */
int ghost_bor_system_count(ghost_bor* h, int* n);
/* Row interchanges of system sys_idx's LU factorization after a solve:
 * ipiv[r] (n entries) is the row swapped with row r at elimination step r,
 * 0-based within the system.  No counterpart in the reference, whose
 * numpy.linalg.solve (bor_solver.py:753) keeps LAPACK zgetrf's ipiv inside:
 * this is that ipiv minus one, for callers that check or reuse the factor. */
int ghost_bor_get_pivots(ghost_bor* h, int sys_idx, int* ipiv);
/* ||A X - B||_F / ||B||_F for every system (bor_solver.py:754-756), out[system count]. */
int ghost_bor_residual(ghost_bor* h, double* out);
/* Mode-system layout of (prob, m): size n, active J_t count nt and first active t / phi nodes. */
int ghost_bor_system_info(ghost_bor* h, int prob, int m, int* n, int* nt, int* t0, int* f0);
int ghost_bor_num_systems(ghost_bor* h, int prob, int* nmodes, int* m_max);
/* Copy a system out (complex, row-major): which = 0 -> A [n x n] and B [n x nrhs]
 * (after fill, before solve); which = 1 -> X [n x nrhs] (after solve). */
int ghost_bor_get_system(ghost_bor* h, int prob, int m, int which, double* out1, double* out2);
/* Zero-pivot count per system from the last factorisation, out[system count]. */
int ghost_bor_info(ghost_bor* h, int* out);
/* Milliseconds of the last fill, solve and far-field phases (HIP events on the
 * handle's stream); -1 for a phase that has not run eagerly since build. */
int ghost_bor_timings(ghost_bor* h, float* ms3);

/* Optional per-launch profile: with on != 0 every later launch is bracketed by
 * HIP events on the handle's stream (and the record is cleared). */
int ghost_bor_set_profile(ghost_bor* h, int on);
/* Number and names of the timed kernel families. */
int ghost_bor_kernel_count(void);
const char* ghost_bor_kernel_name(int kind);
/* Per family since set_profile: summed ms, launch count, algorithmic real flops. */
int ghost_bor_kernel_times(ghost_bor* h, double* ms, int* count, double* flops);

/* Adjoint of a real far-field objective J w.r.t. the radial coordinate of
 * every generatrix node (no counterpart in the reference, whose solver has no
 * gradient; it is the "adjoint solve + dZ/dp contraction" entry of the
 * device-resident interface).  acoef: complex, per problem n_rhs entries in
 * column order 2 theta_index + (0 VV, 1 HH), with dJ = Re sum_c acoef_c dF_c
 * and F_c the mode-summed far-field amplitude (-j k eta0 / 4 pi included).
 * grad: per problem Nn doubles, dJ/d rho_node (0 for nodes within 2 step of
 * the axis).  Method: one transposed solve lambda = A^-T (a' dF/dX) through
 * the LU factors of the last solve, Q = [-Lambda X^T | Lambda], and FP64
 * central differences (half-width step in metres; 0 = 1e-6 of each
 * problem's shortest element) of the fill, right-hand side
 * and far-field projection with the quadrature rules of the build frozen,
 * moving the test and the source geometry separately, one residue class of
 * nodes mod 3 at a time (12 batched fills).  Needs a GHOST_FP64 handle after
 * fill + solve; afterwards the systems hold a perturbed fill. */
int ghost_bor_adjoint(ghost_bor* h, const double* acoef, double step, double* grad);
/* The same gradient with the fill differentiated analytically (no counterpart
 * in the reference either): after the same transposed solve and Q, the
 * kernels contract Q with dA/d rho, dB/d rho and the far-field projection's
 * d/d rho carried by forward-mode dual numbers through the point positions,
 * weights, tangents, azimuth samples and the near pairs' own sinh-graded
 * rules, so the result is the exact derivative of the discrete operator
 * (no step, no refill; the systems keep the forward factors).  grad is 0 for
 * nodes on the axis.  Handles as ghost_bor_adjoint, or GHOST_MIXED: lambda
 * then comes from the FP32 factors under one FP64 refinement sweep. */
int ghost_bor_adjoint_analytic(ghost_bor* h, const double* acoef, double* grad);
/* Forward RCS and its gradient in one graph launch (no counterpart in the
 * reference).  set_objective: J = sum_c w_c sigma_dBsm(c) with real weights
 * per problem column (column order as acoef).  capture_adjoint: records fill,
 * solve with refine_iters fixed FP64 refinement sweeps (>= 0), far field, the
 * objective's a_c = w_c (20 / ln 10) conj(F_c) / |F_c|^2 formed on the device,
 * the transposed solve, Q and the analytic gradient into the handle's graph
 * (replayed by ghost_bor_launch).  get_gradient: dJ/d rho of every node of
 * every problem (problem order) after the last analytic adjoint, eager or
 * captured.  FP64 or GHOST_MIXED handles with mode pairing off. */
/* New node coordinates [Nn][2] for built problem prob (same node count, the
 * same ends on the axis; no counterpart in the reference): the point, node
 * and element tables are recomputed in place, so one built handle, and a
 * graph captured on it, serve a family of shapes without reallocating.  The
 * build's quadrature rules stay (n_xi, the near pairs' tail orders).  Plain
 * conductor problems only. */
int ghost_bor_set_nodes(ghost_bor* h, int prob, const double* nodes);
int ghost_bor_set_objective(ghost_bor* h, const double* weights);
int ghost_bor_capture_adjoint(ghost_bor* h, int refine_iters);
int ghost_bor_get_gradient(ghost_bor* h, double* grad);

/* Right-hand sides only (Bessel table and excitation of every problem): the
 * device side of rhs_mode / rhs_mfie_mode (bor_solver.py:437-461, 625-645). */
int ghost_bor_fill_rhs(ghost_bor* h);
/* Put a caller's solution X (n x n_rhs, masked rows, complex row-major) of
 * mode m where ghost_bor_farfield reads it: the device side of farfield_mode
 * (bor_solver.py:659-708) for solutions computed elsewhere. */
int ghost_bor_set_solution(ghost_bor* h, int prob, int m, const double* X);

/* Caller-filled dense complex systems solved in the same batched LU as the
 * mode systems: the device side of the reference's generic mode loop
 * (_mode_sweep, bor_solver.py:711-801), whose np.linalg.solve (line 753) this
 * replaces for closures that assemble their own systems.  add_dense before
 * build; set_dense (A: n x n, B: n x n_rhs, complex row-major) after build and
 * before each solve (the LU works in place); get_dense (X: n x n_rhs) after
 * the solve.  Dense systems have no fill and no far field. */
int ghost_bor_add_dense(ghost_bor* h, int n, int nrhs, int* sys_idx);
int ghost_bor_set_dense(ghost_bor* h, int sys_idx, const double* A, const double* B);
int ghost_bor_get_dense(ghost_bor* h, int sys_idx, double* X);

/* Measurement utility (no counterpart in the reference): one plain device
 * copy of `bytes` bytes (read once, written once) on its own stream, *ms its
 * duration; the known byte count calibrates HBM byte counters. */
int ghost_copy_calibration(double bytes, float* ms);

/* Host-only query (no device call): wavefronts per far-fill work-group for a
 * problem of this precision, m_max and azimuth grid sizes, as chosen by the
 * LDS budget (tables + per-wavefront order stash within 160 KiB); 0 means the
 * combination does not fit and ghost_bor_add_problem rejects it.  No
 * counterpart in the reference (a tile-size query of the device interface). */
int ghost_far_fill_waves(int precision, int m_max, int n_xi_efie, int n_xi_bracket);

#ifdef __cplusplus
}
#endif

#endif /* GHOST_BOR_H */
