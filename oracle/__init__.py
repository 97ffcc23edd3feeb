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

How each function is pinned (tests/test_oracle_*.py):
  fem.prepare_mesh / gen_fwdmat : closed-form P1 element integrals, sum(M) =
      volume, K_stiff @ 1 = 0, 1'K1 = sum(int mua) + R*(true + extra face area),
      face counts of the Kuhn slab, Robin constant R for n = 1.4.
  fsai.fsaip_gen               : exact FSAI identities (G A G')_ii = 1 and
      (G A)_ij = 0 on the pattern; diagonal A gives 1/sqrt(a_ii); brute-force
      dense solve per row; pattern arithmetic vs a brute-force restatement.
  solver.bicgstab_fsai         : converged x vs a direct sparse LU solve.
  transforms.*                 : td2fc of a Gaussian vs its closed form; the
      fc2tg = (Nt/(Nt-1)) sum_t g(t) fc2td(F)(t) identity; fc2td o td2fc on a
      band-limited periodic signal.
  jacobian.j_complex           : central finite differences of meas_flu in
      nodal mua (exact-derivative mode, kappa fixed) for the nodal-product
      approximation's limit; reciprocity.
  recon.tik_inv / reshape_J    : push-through identity, numpy.linalg.solve.
Parity unpinned (no reference run exists: MATLAB is absent): roundoff order of
MATLAB's sparse() summation, pointLocation tie-breaking, BiCGSTAB iteration
trajectories.
"""
