"""FP64 CPU oracle for DOGPUP's frequency-domain hot path.

TEST INFRASTRUCTURE ONLY.  This package restates the reference's MATLAB
algorithms (``DOGPUP/dMesh.m``, ``dOptode.m``, ``td2fc.m``, ``fc2td.m``,
``fc2tg.m``, ``id_thresh.m``, ``reconstruction/*.m``) and the behaviour of its
CUDA kernels (``cuda_lib/mex_sources/*_CUDA.cu``) in NumPy/SciPy, in FP64, with
the reference's own structure (quadrature loops, lockstep BiCGSTAB, per-row
Gaussian elimination).  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only to check results or
as the timed CPU baseline; the product (``dogpup_amd``) never imports it.

Indices are 0-based throughout (the reference is 1-based in MATLAB and 0-based
in its CSR arrays).  Arrays keep the reference's axis order: fluence is
NN x NF x NS, boundary data NM x NF, the Jacobian NM x NF x NV.

How each function is pinned (tests/test_oracle_*.py, tests/test_golden.py):
  fem.prepare_mesh / gen_fwdmat : closed-form P1 element integrals, sum(M) =
      volume, K_stiff @ 1 = 0, 1'K1 = sum(int mua) + R*(boundary area), face
      counts of the Kuhn slab and the reference's face list, the Robin
      constant R for n = 1.4.
  fsai.pattern_rows_cols       : a literal list-operation restatement of the
      window (largest entry + the N-1 entries listed before it), and what it
      means per row (largest + N-1 smallest when a row has more than N).
  fsai.fsaip_gen / _batched    : (G A G')_ii = 1 and (G A)_ij = 0 on the
      pattern; diagonal A gives 1/sqrt(a_ii); numpy.linalg.solve per row;
      the batched form equals the per-row form.
  solver.bicgstab_fsai / _core : converged x vs a direct sparse LU solve.
  transforms.*                 : td2fc of a Gaussian vs its closed form;
      fc2tg = (Nt/(Nt-1)) sum_t g(t) fc2td(F)(t); fc2td o td2fc on a
      band-limited periodic signal; DC halving and dim handling; id_thresh's
      falling edge.
  optode.snap_pos / snap2mesh  : placements and nodal weights on a lattice
      mesh (weight 1 at a node, 1/2 + 1/2 at an edge midpoint, inward
      normals, the 1/min(musp) source shift onto the node below).
  optode.meas_flu              : reciprocity d . phi_s = q_s . phiA_d.
  jacobian.j_exact / j_complex : central differences of meas_flu in nodal
      mua for one link and for a 2 x 2 link table; j_complex within 15 % of
      the exact derivative, which the swapped link table fails; j_tg =
      fc2tg of j_complex.
  grid.*                       : Kuhn cells tile and are Delaunay, linear
      reproduction, lattice identities, hand-computed grid_to_mesh rows.
  recon.tik_inv / reshape_J    : push-through identity, orthogonal-column
      closed form, reshape_J limits.
  lm.*                         : the spot prior's shape (FWHM), convergence
      to the truth on a log-linear model whose Jacobian is exact, the
      step-back branch taken twice.
  cpu (OpenMP C++)             : per-row elimination and the NumPy oracle's
      BiCGSTAB, and the direct solve.
  everything on tests/golden/tiny_slab.npz (committed inputs and outputs).
Parity unpinned (no reference run exists: MATLAB is absent): roundoff order of
MATLAB's sparse() summation, pointLocation's choice between the elements that
share a face, BiCGSTAB iteration trajectories.
"""
