/*
 * dogpup.h -- C ABI of libdogpup_hip.so, the MI355X (gfx950) implementation of
 * DOGPUP's frequency-domain diffuse-optics hot path.
 *
 * Conventions
 *  - Every entry point returns an int status: DOGPUP_OK (0) or a negative
 *    DOGPUP_ERR_*; dogpup_last_error() gives the message of the last failure on
 *    the calling thread.  No C++ exception crosses this boundary.
 *  - Threads: the library keeps no mutable global state other than the
 *    per-device scratch array of dogpup_fc2td_field_noalloc (whose address
 *    lookup is locked).  A handle (solver,
 *    locator) is not locked: use one handle per thread, or serialise the
 *    calls on a handle; different handles may run on different threads at
 *    once (graph capture is thread-local to the capturing thread).
 *  - CSR arrays are int32 and 0-based, as the reference's sparse_csr produces
 *    (DOGPUP/dMesh.m:1388-1414).
 *  - "dev" pointers are device (HBM) pointers, "host" pointers are host memory.
 *    `stream` is a hipStream_t (NULL = default stream); device entry points
 *    are asynchronous on it unless stated otherwise.
 *  - Complex data are interleaved (re, im) doubles or floats unless an entry
 *    point mirrors a reference launcher with planar re/im arrays.
 *  - Reference layout ("MATLAB layout") of a multi-RHS block: NN x NS x NF
 *    column-major, element (row, col, freq) at freq*NS*NN + col*NN + row
 *    (smv_CUDA.cu:32,41).
 *  - Solver layout of a multi-RHS block: groups of NSP columns, one group per
 *    (frequency, column block): element (g, node, c) at (g*NN + node)*NSP + c,
 *    g = freq*NCB + col/NSP, c = col % NSP, NSP in {8,16,32,64}.
 */
#ifndef DOGPUP_H
#define DOGPUP_H

#ifdef __cplusplus
extern "C" {
#endif

#define DOGPUP_OK 0
#define DOGPUP_ERR_ARG -1
#define DOGPUP_ERR_HIP -2
#define DOGPUP_ERR_UNSUPPORTED -3
#define DOGPUP_ERR_NOMEM -4

/* per-RHS solver status */
#define DOGPUP_RHS_CONVERGED 0
#define DOGPUP_RHS_MAXIT 1
#define DOGPUP_RHS_BREAKDOWN 2
#define DOGPUP_RHS_ZERO 3 /* zero right-hand side: x = 0 */

/* solver precision / structure modes */
#define DOGPUP_MODE_C128 0          /* complex128 fused BiCGSTAB, per-RHS convergence */
#define DOGPUP_MODE_C64_REFINE 1    /* complex64 fused BiCGSTAB inside FP64 iterative refinement */
#define DOGPUP_MODE_C128_PLAIN 2    /* complex128, reference structure: unfused passes, lockstep */

/* Krylov method of the fused forms (DOGPUP_MODE_C128, DOGPUP_MODE_C64_REFINE) */
#define DOGPUP_METHOD_BICGSTAB 0    /* the reference's BiCGSTAB (dMesh.m:1247-1348); the plain form's only one */
#define DOGPUP_METHOD_COCG 1        /* conjugate orthogonal CG: CG with the unconjugated form; the fused forms' default */

const char* dogpup_last_error(void);
int dogpup_version(void);
int dogpup_device_count(int* count);

/* ---------------------------------------------------------------------------
 * Reference-compatible launchers: device pointers, planar FP64 re/im arrays,
 * reference layout.  Same argument order and meaning as the reference's
 * launchers, plus a trailing stream.
 * ------------------------------------------------------------------------- */

/* Replaces smv_call (DOGPUP/cuda_lib/mex_sources/smvMex.hpp:4-15,
 * smv_CUDA.cu:59-82): y(:,c,k) = (Re A + i*fAxis[k]*Im A) x(:,c,k). */
int dogpup_smv_call(double* d_out_real, double* d_out_imag, const double* d_x_real,
                    const double* d_x_imag, const double* d_fAxis, const double* d_valA_real,
                    const double* d_valA_imag, const int* d_rPtr, const int* d_cPtr, int nrows,
                    int ncols, int nfreqs, void* stream);

/* Replaces FSAImv_call (FSAImvMex.hpp:4-15, FSAImv_CUDA.cu:58-81):
 * y(:,c,k) = G_k x(:,c,k), G_k values at valG[nnz*k + i]. */
int dogpup_fsaimv_call(double* d_out_real, double* d_out_imag, const double* d_x_real,
                       const double* d_x_imag, const double* d_valG_real,
                       const double* d_valG_imag, const int* d_rPtr, const int* d_cPtr, int nnz,
                       int nrows, int ncols, int nfreqs, void* stream);

/* Replaces FSAIP_call (FSAIP.hpp:4-13, FSAIP_CUDA.cu:162-180): FSAI factor rows
 * of A_k = Re A + i*k*Im A (A given as its lower triangle tril(A) in CSR) for
 * k = 0..nfreqs-1 on the pattern (rPtrG, cPtrG) (ascending columns, diagonal
 * last).  Output NF x nnzG planar, frequency fastest: out[k + nfreqs*j].
 * Rows with more than 32 pattern entries return DOGPUP_ERR_UNSUPPORTED. */
int dogpup_fsaip_call(double* d_out_real, double* d_out_imag, const double* d_valA_real,
                      const double* d_valA_imag, const int* d_rPtrA, const int* d_cPtrA,
                      const int* d_rPtrG, const int* d_cPtrG, int nfreqs, int nrows,
                      void* stream);

/* ---------------------------------------------------------------------------
 * Mex-level entry points: host arrays in the interleaved-complex MATLAB
 * layout (R2018a complex API), synchronous.  The MATLAB gateways in
 * dogpup_amd/mex/ forward their mxArrays here (gpuArray is CUDA-only).  Each
 * CSR structure is checked on the host first (row pointer from 0 to nnz,
 * never decreasing; columns in [0, nrows)): DOGPUP_ERR_ARG otherwise.
 * ------------------------------------------------------------------------- */

/* y = smv_mex(x, f, A, rPtrA, cPtrA)  (smv_mexgateway.cpp:8-100).
 * x, y: 2*nrows*ncols*nfreqs doubles; f: nfreqs doubles; A: 2*nnz doubles. */
int dogpup_smv_mex(double* h_y, const double* h_x, const double* h_f, const double* h_A,
                   const int* h_rPtrA, const int* h_cPtrA, int nrows, int ncols, int nfreqs,
                   int nnz);

/* y = FSAImv_mex(x, G, rPtrG, cPtrG)  (FSAImv_mexgateway.cpp:8-100).
 * G: nnzG x nfreqs complex, column-major (2*nnzG*nfreqs doubles). */
int dogpup_fsaimv_mex(double* h_y, const double* h_x, const double* h_G, const int* h_rPtrG,
                      const int* h_cPtrG, int nrows, int ncols, int nfreqs, int nnzG);

/* valG = FSAIP_gen_mex(valA, rPtrA, cPtrA, rPtrG, cPtrG, nfreqs)
 * (FSAIP_mexgateway.cpp:8-85).  valA: tril(A) CSR values (2*nnzA doubles);
 * output NF x nnzG complex column-major (2*nfreqs*nnzG doubles). */
int dogpup_fsaip_gen_mex(double* h_valG, const double* h_valA, const int* h_rPtrA,
                         const int* h_cPtrA, const int* h_rPtrG, const int* h_cPtrG,
                         int nrows, int nnzA, int nnzG, int nfreqs);

/* ---------------------------------------------------------------------------
 * FEM assembly (replaces gen_fwdmat, dMesh.m:1033-1194, and the formation of
 * the full complex system in fluGPU, dMesh.m:1241-1242).
 *
 * Produces the values of the full symmetric A on a CSR pattern prepared by
 * the host: valA_re = K (stiffness(kappa) + absorption mass(mua) + Robin
 * term), valA_im = scale_im * M (scale_im = df/c, unit frequency).
 * slot_ptr (nnz+1) / slot_src (contributions) list, per CSR slot, the element
 * (code e*16 + a*4 + b) and Robin-face (code NE*16 + f*9 + a*3 + b)
 * contributions that sum into it.  work: NE*17 doubles of device scratch.
 * ------------------------------------------------------------------------- */
int dogpup_assemble(int nn, int ne, int nbf, int nnz, const double* d_node, const int* d_elem,
                    const int* d_bface, const double* d_mua, const double* d_kappa,
                    const double* d_R, const int* d_slot_ptr, const int* d_slot_src,
                    double scale_im, double* d_work, double* d_valA_re, double* d_valA_im,
                    void* stream);

/* ---------------------------------------------------------------------------
 * Batched FSAI-BiCGSTAB over every (frequency, column) right-hand side
 * (replaces gpuBicstab_FSAIP, dMesh.m:1247-1348, and FSAIP_gen's factor call,
 * dMesh.m:1384).  A solver handle keeps the pattern, the matrix, the FSAI
 * factors of every frequency and all workspace resident in HBM.
 * ------------------------------------------------------------------------- */
typedef struct dogpup_solver* dogpup_solver_t;

/* d_rptrA/d_cptrA: full symmetric CSR of A (sorted columns).  d_rptrG/d_cptrG:
 * FSAI pattern (lower, ascending, diagonal last), both in the caller's node
 * order; the FSAI factors are computed on these, as the reference does.
 * h_order (host, nn entries, or NULL for the identity) is the node order the
 * iterations run in: solver row i holds caller node h_order[i] (a
 * space-filling-curve order keeps each row tile's stencil in cache).  The
 * solve takes and returns vectors in the caller's order.  ncols = columns
 * per frequency (sources); nsp = group width (8/16/32/64).  The handle keeps
 * the caller's pattern pointers (caller keeps them alive). */
int dogpup_solver_create(dogpup_solver_t* out, int nn, int nfreq, int ncols, int nsp, int mode,
                         int nnzA, const int* d_rptrA, const int* d_cptrA, int nnzG,
                         const int* d_rptrG, const int* d_cptrG, const int* h_order);
/* Batch of independent meshes in one handle (one launch per pass over all
 * of them).  The caller concatenates the meshes' patterns into one
 * block-diagonal CSR: mesh i owns rows [row0[i], row0[i+1]) (row0[0] = 0,
 * nn = row0[nmesh]) and its column indices stay inside that range, in A and in
 * the FSAI pattern G alike; h_order (optional) must map each mesh's rows
 * into the same mesh.  Mesh i has nfreq[i] frequencies and ncols[i] RHS
 * columns.  dogpup_solver_set_matrix then takes the concatenated values and
 * sum(nfreq) frequency scales, mesh after mesh; b and x hold mesh after mesh
 * the single-mesh solver layout of each (nfreq[i] * ceil(ncols[i]/nsp)
 * groups of row0[i+1]-row0[i] rows); the per-RHS outputs hold mesh after mesh
 * freq*ncols[i] + col.  A single mesh through dogpup_solver_create is the
 * case nmesh = 1.  No counterpart in the reference, which solves one mesh per
 * call. */
int dogpup_batch_create(dogpup_solver_t* out, int nmesh, const int* h_row0, const int* h_nfreq,
                        const int* h_ncols, int nsp, int mode, int nnzA, const int* d_rptrA, const int* d_cptrA,
                        int nnzG, const int* d_rptrG, const int* d_cptrG, const int* h_order);

/* Kernel shapes of the following solves (0 keeps the current value): row
 * tile of the SpMV and vector passes (a multiple of 64), x gathers in flight
 * per row in the A and G SpMVs (4, 8 or 16), captured iterations replayed
 * between convergence polls.  Chosen by dogpup_amd.autotune. */
int dogpup_solver_set_tuning(dogpup_solver_t s, int tile_rows, int unroll_a, int unroll_g, int graph_k);

/* Level-0 spacing of the frequency continuation (see
 * dogpup_solver_set_freq_continuation): every stride-th frequency and the
 * last are solved from x0 = 0, each later level halves the spacing.  stride
 * a power of two >= 2; 0 restores the default (16 with Galerkin guesses, 8
 * with Lagrange guesses).  Chosen by dogpup_amd.autotune with the kernel
 * shapes. */
int dogpup_solver_set_cont_stride(dogpup_solver_t s, int stride);

int dogpup_solver_destroy(dogpup_solver_t s);

/* Set A (FP64 values on the handle's pattern; frequency k uses
 * valA_re + i*fscale[k]*valA_im) and compute the FSAI factors of every
 * frequency (FP64), G^T, and their complex64 copies when the mode needs them.
 * d_fscale: nfreq doubles (the reference uses k = 0..NF-1).  A call with the
 * previous call's fscale (new optical properties at the same frequencies)
 * keeps the handle's captured iteration graphs and continuation schedule;
 * other scales rebuild them at the next solve.  After a failed call the
 * handle has no matrix until a call succeeds. */
int dogpup_solver_set_matrix(dogpup_solver_t s, const double* d_valA_re, const double* d_valA_im,
                             const double* d_fscale, void* stream);

/* Solve A_k x = b for every RHS.  b, x: complex128 in solver layout (device).
 * tol: relative residual ||b - A x|| / ||b|| per RHS (FP64 true residual in
 * C64_REFINE mode); maxit: BiCGSTAB iterations (all sweeps together in
 * C64_REFINE).  Outputs (host, nfreq*ncols each, index freq*ncols + col):
 * iterations, final relative residual, status.  Synchronous.  d_x may be
 * d_b (in place).  With the node order the identity the solve iterates in
 * d_x itself (no row permutation), and a later solve into another buffer
 * recaptures the FP64 fused form's iteration graph. */
int dogpup_solver_solve(dogpup_solver_t s, const double* d_b, double* d_x, double tol, int maxit,
                        int* h_iters, double* h_relres, int* h_status, void* stream);

/* Krylov method for the following solves (DOGPUP_METHOD_*).  Both converge
 * to the same x on the same tolerance; A_k is complex symmetric and
 * M^-1 = G_k^T G_k is symmetric, which is what COCG needs.  COCG applies three
 * operators per iteration (A, G, G^T) and moves 12 vectors, BiCGSTAB seven
 * and 30.  No counterpart in the reference, which hard-codes BiCGSTAB
 * (dMesh.m:1247-1348); the reference-structure mode is BiCGSTAB only. */
int dogpup_solver_set_method(dogpup_solver_t s, int method);

/* Per-RHS tolerances for the following solves: h_tol[o] is the relative
 * residual at which right-hand side o (output order: per mesh,
 * freq * ncols + col) counts as converged; 0 keeps the tol passed to
 * dogpup_solver_solve.  n = 0 or h_tol = NULL clears them.  The fused and
 * refinement forms only (the reference-structure form iterates in lockstep to
 * one tolerance).  It lets a batched forward + adjoint solve run each half to
 * the accuracy its result needs, as separate flu_solve and adj_flu_solve calls
 * (dMesh.m:106-202) would with two tolerances; no counterpart in the
 * reference, whose gpuBicstab_FSAIP hard-codes one (dMesh.m:1243). */
int dogpup_solver_set_rhs_tol(dogpup_solver_t s, const double* h_tol, int n);

/* Initial guesses of the following solves (DOGPUP_MODE_C64_REFINE): the
 * first FP64 iterate of right-hand side (f, col) is
 * sum_q w_q * x(h_src[f*nterm + q], col), w_q the complex weight
 * (h_w[2*(f*nterm + q)], h_w[2*(f*nterm + q) + 1]), where x(f', col) is the
 * converged solution at frequency f' (same mesh, same column) of the same
 * solve, or, for h_src = f itself, what d_x holds for (f, col) on entry to
 * dogpup_solver_solve (a warm start from an earlier solve; d_x must then be
 * finite).  h_src = -1 ends a frequency's terms; frequencies are global over
 * the handle's meshes, as in d_fscale.  The solve runs in levels: a
 * frequency without terms naming other frequencies is in level 0, any other
 * one level after the latest frequency it names, and each level is solved
 * to convergence before the next starts.  Convergence is on the same
 * ||b - A x|| / ||b|| <= tol.  nterm = 0 clears the guesses (x0 = 0); 1..8
 * terms.  No counterpart in the reference, which starts every RHS from
 * x0 = 0 (dMesh.m:1283). */
int dogpup_solver_set_guess(dogpup_solver_t s, int nterm, const int* h_src, const double* h_w);

/* Frequency continuation (on by default in DOGPUP_MODE_C64_REFINE, with
 * Galerkin guesses and a flat spectrum; the environment variable
 * DOGPUP_FREQ_CONT=0 turns the default off): guesses built from the
 * frequency scales of dogpup_solver_set_matrix.  The frequencies of a mesh
 * are solved in levels, each level to convergence before the next: level 0
 * is every stride-th frequency and the last, and each later level halves the
 * spacing (stride 16 for Galerkin guesses, 8 for Lagrange guesses).  A
 * later-level frequency starts from a combination of the converged solutions
 * of earlier levels nearest to it in the frequency scale:
 *   on = DOGPUP_CONT_GALERKIN (1): the six nearest, combined per column as
 *     the Galerkin solution of its system in their span (the M-Gram matrix
 *     of the solutions and their dots with the right-hand side give that
 *     small system exactly, see dogpup_amd/csrc/solver.hip);
 *   on = DOGPUP_CONT_LAGRANGE (2): the four nearest, combined by Lagrange
 *     interpolation in the frequency scale.
 * h_spectrum (NULL, or nf_total complex values): the right-hand side of
 * frequency f is h_spectrum[f] times a vector that does not depend on f, as
 * for point sources, whose right-hand sides carry the source spectrum s_fpsf
 * (dMesh.m:128,196); the Lagrange interpolation then runs on x / spectrum,
 * which is smooth in the frequency where x is not, and the Galerkin system
 * uses it to express K x_j through b_j.  Convergence is the same
 * ||b - A x|| / ||b|| <= tol per RHS whatever the guess.  on = 0 solves every
 * RHS from x0 = 0, as the reference does; either call clears guesses set with
 * dogpup_solver_set_guess. */
#define DOGPUP_CONT_GALERKIN 1
#define DOGPUP_CONT_LAGRANGE 2
int dogpup_solver_set_freq_continuation(dogpup_solver_t s, int on, const double* h_spectrum);

/* Counters of the last solve: number of vector-kernel passes weighted by the
 * columns they moved, for roofline accounting.  h_out[0] = BiCGSTAB
 * iterations (max), [1] = sum over groups of active iterations (c64 or
 * c128), [2] = FP64 refinement sweeps, [3] = inner sweeps. */
int dogpup_solver_stats(dogpup_solver_t s, double* h_out, int n);

/* Per-launch timing of the next solves (HIP events on the solve's stream).
 * dogpup_solver_profile fills h_out[3*c + {0,1,2}] = {milliseconds,
 * algorithmic bytes, launches} for kernel class c: 0 = G/G^T SpMV, 1 = A
 * SpMV, 2 = BiCGSTAB vector passes, 3 = FP64 refinement passes. */
int dogpup_solver_set_profile(dogpup_solver_t s, int on);
int dogpup_solver_profile(dogpup_solver_t s, double* h_out, int n);

/* The handle's FSAI factors in the caller's pattern order (FP64, [freq][nnzG]
 * interleaved complex, 2*nfreq*nnzG doubles) into d_out.  When
 * dogpup_solver_set_matrix wrote the factors straight into the solver's
 * layout (every pattern row of at most 8 entries), the first call after it
 * computes this FP64 copy (same kernel arithmetic), then keeps it until the
 * next dogpup_solver_set_matrix. */
int dogpup_solver_get_fsai(dogpup_solver_t s, double* d_out, void* stream);

/* Layout transforms between MATLAB layout (NN x NF x NS column-major,
 * interleaved complex128) and solver layout. */
int dogpup_layout_to_solver(const double* d_src, double* d_dst, int nn, int nfreq, int ncols,
                            int nsp, void* stream);
int dogpup_layout_from_solver(const double* d_src, double* d_dst, int nn, int nfreq, int ncols,
                              int nsp, void* stream);

/* Point-source right-hand sides in solver layout (flu_solve / adj_flu_solve
 * source vectors, dMesh.m:124-129,191-196): b(g,node,c) = sum_q w[c][q] *
 * spectrum[freq] for node = nodes[c][q]; columns >= ncols are zero. */
int dogpup_point_sources(double* d_b, int nn, int nfreq, int ncols, int nsp, const int* d_nodes,
                         const double* d_w, const double* d_spectrum, void* stream);

/* ---------------------------------------------------------------------------
 * Fourier-series transforms along the leading axis of a block with `ncol`
 * contiguous columns (the reference permutes the transformed dimension to
 * the front the same way).  f: angular frequencies (rad/s), t: time axis (s),
 * device arrays.  Synchronous on `stream`.
 * ------------------------------------------------------------------------- */

/* Replaces td2fc (DOGPUP/td2fc.m:1-39): out(k,col) = dt/P sum_t in(t,col)
 * exp(-i f_k t), P = t(end).  in: NT x ncol real (in_complex = 0) or complex;
 * out: NF x ncol complex. */
int dogpup_td2fc(const double* d_in, double* d_out, const double* d_f, const double* d_t, int nt,
                 int nf, int ncol, int in_complex, void* stream);

/* Replaces fc2td (fc2td.m:1-37): out(t,col) = 2 Re sum_k w_k in(k,col)
 * exp(i f_k t), w_0 = 1/2.  in: NF x ncol complex; out: NT x ncol real. */
int dogpup_fc2td(const double* d_in, double* d_out, const double* d_f, const double* d_t, int nf,
                 int nt, int ncol, void* stream);

/* fc2td (fc2td.m:24-31) of a whole field on FP64 MFMA with FP32 outputs:
 * out[t][e] = 2 Re sum_k w_k F[k][e] exp(i f_k t_t), w_0 = 1/2, for every
 * column e < ncol of a complex128 block laid out [k][e] (a solver-layout
 * field: ncol = groups-per-frequency * nn * nsp, so out is the time-domain
 * field in the same [t][group][node][column] order).  This is the
 * reference's fc2td(phi, f, t, 2) over nodal fluence (NN x NF x NS ->
 * NN x NT x NS) with the layout of the solver.  Products and sums are FP64
 * (a bin of a curve far from its source is a near-total cancellation of the
 * 2*nf terms); only the stored value is FP32.  nf <= 128.  Synchronous. */
int dogpup_fc2td_field(const double* d_field, float* d_out, const double* d_f, const double* d_t, int nf,
                       int nt, long long ncol, void* stream);

/* Allocation-free dogpup_fc2td_field: same arguments, result and limits
 * (nf <= 128), asynchronous on `stream`, and no device allocation or
 * host synchronisation.  No counterpart in the reference: it is for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops).
 * The per-pass DFT table lives in a static device array of 133,120 bytes
 * (16,640 doubles) that every GPU the library runs on holds until the
 * process ends; the first call on a device resolves that array's address
 * (make it once before capturing a graph).  All calls on one GPU share the
 * array: calls that overlap on one GPU (other streams or threads) race. */
int dogpup_fc2td_field_noalloc(const double* d_field, float* d_out, const double* d_f, const double* d_t, int nf,
                               int nt, long long ncol, void* stream);

/* Replaces fc2tg (fc2tg.m:1-38): out(g,col) = nt_factor * 2 Re sum_k w_k
 * conj(gates(g,k)) in(k,col).  gates: NG x NF complex (row-major);
 * nt_factor = Nt (number of time samples). */
int dogpup_fc2tg(const double* d_in, double* d_out, const double* d_gates, int nf, int ng, int ncol,
                 double nt_factor, void* stream);

/* Replaces the detector sampling of meas_flu (dMesh.m:144-171):
 * out(m,k) = sum_q dw[d_m][q] phi(dnodes[d_m][q], k, s_m).  phi: solver
 * layout (NS columns); dnodes/dw: ND x 4; link: NM x 2 int32 (source,
 * detector), 0-based; out: NM x NF complex128, m fastest. */
int dogpup_meas(const double* d_phi, double* d_out, const int* d_dnodes, const double* d_dw,
                const int* d_link, int nn, int nf, int ncols, int nsp, int nm, int nd, void* stream);

/* Replaces J_complex followed by fc2tg (dMesh.m:206-247, wghtd_LM_recon.m:
 * 58-61) on the nodal basis (m2g = I):
 * out[v][g][m] = scale(m,g) * sum_k Re(coef[g][k] phi(v,k,s_m) phi(v,k,d_m)),
 * coef[g][k] = -2 Nt w_k conj(gate(g,k)) dxyz^3 / s_fpsf(k) (complex128,
 * NG x NF).  d_field: forward and adjoint fluence of one batched solve,
 * complex128 solver layout (ncols columns per frequency, group width nsp);
 * link: NM x 2 int32 = (source column, detector column) in that field;
 * scale: NG x NM float (m fastest) or NULL; out: float, NM x NG x NV
 * column-major (node v = 0..nv-1).  Sums in FP64 (only out is FP32).
 * Limits: links within 32 source and 32 detector columns without repeated
 * pairs take the MFMA form (nf <= 128, any ng); otherwise ng <= 16 and
 * (nf*ncols + ng*nf)*16 <= 160 KiB. */
int dogpup_jacobian_tg(const double* d_field, const double* d_coef, const int* d_link,
                       const float* d_scale, float* d_out, int nn, int nf, int ng, int ncols, int nsp,
                       int nm, int nv, void* stream);

/* Vector-Jacobian product of the time-gated absorption Jacobian, without
 * forming J (the backward of mua -> time-gated data, dMesh.m:206-247 with
 * fc2tg): g = J_TG^T ybar.
 * d_field: forward + adjoint fluence of one solve (complex128 solver layout,
 * ncols columns, sources in columns 0..ns-1); d_D: NF x NM complex128,
 * D[k][m] = sum_g ybar[m][g] * 2 Nt w_k conj(gate[g][k]) / s_fpsf[k];
 * link: NM x 2 (source column, detector column); link_ptr/link_idx: links
 * grouped by source (CSR over sources).  exact = 0: the reference's
 * nodal-product Jacobian (scale dxyz3 = dxyz^3); exact = 1: the exact FEM
 * derivative at fixed kappa, element-local int N_v N_i N_j (needs node,
 * elem).  d_E: NN*NF*NS complex128 workspace; d_g: NN doubles (output). */
int dogpup_jacobian_vjp(const double* d_field, const double* d_D, const int* d_link,
                        const int* d_link_ptr, const int* d_link_idx, const double* d_node,
                        const int* d_elem, double* d_E, double* d_g, int nn, int ne, int nf, int ncols,
                        int nsp, int ns, int nm, int exact, double dxyz3, void* stream);

/* ---------------------------------------------------------------------------
 * Point location and grid <-> mesh interpolation.  These replace the host
 * triangulation/pointLocation/delaunay calls of mesh2grid (dMesh.m:286-380),
 * mesh2mesh (:383-411), mesh_slice/mesh_slice3 (:414-517) and the optode
 * location of snap2mesh (dOptode.m:77-97).  Device pointers; node NN x 3
 * float64 and elem NE x 4 int32 (0-based), row-major.
 * ------------------------------------------------------------------------- */

typedef struct dogpup_locator* dogpup_locator_t;

/* Bins the tetrahedra of a mesh (a copy of node/elem is kept). */
int dogpup_locator_create(dogpup_locator_t* out, const double* d_node, int nn, const int* d_elem, int ne,
                          void* stream);
int dogpup_locator_destroy(dogpup_locator_t h);
/* h_out[0..2] bins per axis, h_out[3] element-bin pairs. */
int dogpup_locator_stats(dogpup_locator_t h, long long* h_out);

/* Containing tetrahedron of each point (pts: npts x 3): eid (-1 outside),
 * barycentric weights w (npts x 4, NaN outside) and, if d_nodes is not NULL,
 * its 4 node ids (npts x 4, -1 outside).  A point is inside when the largest,
 * over tetrahedra, of the smallest weight is >= -tol; ties go to the smaller
 * element id. */
int dogpup_locator_find(dogpup_locator_t h, const double* d_pts, long long npts, double tol, int* d_eid,
                        double* d_w, int* d_nodes, void* stream);

/* Grid -> mesh interpolation weights of mesh2grid (dMesh.m:331-373): each
 * node is located in the Kuhn split of the voxel grid (origin, spacing, dims
 * per axis; voxel v = ix + nx*(iy + ny*iz)); vertices outside the mesh
 * (inmesh[v] == 0) are dropped and the rest renormalised to sum 1 (a zero
 * sum is left unscaled, as the reference does); a node with no in-mesh
 * vertex takes its nearest in-mesh voxel with weight 1.  Output idx/w:
 * NN x 4 (unused slots -1 and 0).  h_status_counts (optional, 2 ints):
 * nodes that took the nearest voxel, nodes outside the grid (an error). */
int dogpup_grid_to_mesh(const double* d_node, int nn, const double* h_origin, const double* h_spacing,
                        const int* h_dims, const unsigned char* d_inmesh, int* d_idx, double* d_w,
                        int* h_status_counts, void* stream);

/* Sparse interpolation with at most 4 entries per row (m2g, g2m, mesh2mesh):
 * out[g][r][c] = sum_q w[r][q] in[g][idx[r][q]][c] over idx >= 0;
 * rows of `width` doubles (2*nsp for a complex128 solver-layout field). */
int dogpup_ell4_apply(const double* d_in, double* d_out, const int* d_idx, const double* d_w,
                      long long nrows_out, long long nrows_in, int width, int ngroups, void* stream);

/* ---------------------------------------------------------------------------
 * Dense algebra of the weighted LM step (reconstruction/tik_inv.m:11-19,
 * reshape_J.m:13-15): normal matrices on FP32 MFMA with FP32 accumulation,
 * FP64 Cholesky of the damped system.  Row-major, device pointers.
 * ------------------------------------------------------------------------- */

/* C = alpha * A * op(B) from FP32 operands: trans_b = 1 -> op(B) = B^T with
 * B N x K (ldb); trans_b = 0 -> B is K x N.  A is M x K (lda); C is M x N
 * (ldc), FP64 when out_f64 else FP32; lower_only writes only C[i][j], j <= i
 * (for J J^T).  compute_f64 = 0: FP32 MFMA, FP32 accumulation; 1: exact FP64
 * products on FP64 MFMA with FP64 accumulation. */
int dogpup_gemm_f32(const float* d_A, const float* d_B, void* d_C, int M, int N, int K, long long lda,
                    long long ldb, long long ldc, float alpha, int trans_b, int out_f64, int lower_only,
                    int compute_f64, void* stream);

/* out[j] = sum_i A[i][j]^2 (FP64), A rows x cols FP32 (tik_inv's alpha). */
int dogpup_col_sumsq(const float* d_A, int rows, long long cols, double* d_out, void* stream);

/* out[j] = sum_i A[i][j] y[i] (A^T y, FP64 accumulation). */
int dogpup_gemv_t(const float* d_A, const double* d_y, int rows, long long cols, double* d_out,
                  void* stream);

/* In-place lower Cholesky of (A + shift*I), A n x n FP64 row-major (lower
 * triangle used).  *h_info = first non-positive pivot (1-based) or 0;
 * returns DOGPUP_ERR_ARG when the matrix is not positive definite.
 * Synchronous. */
int dogpup_chol_factor(double* d_A, int n, double shift, int* h_info, void* stream);

/* B <- (L L^T)^-1 B, B n x nrhs FP64 row-major, L from dogpup_chol_factor. */
int dogpup_chol_solve(const double* d_L, double* d_B, int n, int nrhs, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* DOGPUP_H */
