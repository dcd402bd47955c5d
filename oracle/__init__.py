"""CPU oracle for the HODLR Gaussian-process path (test infrastructure only).

This package is a plain numpy/scipy FP64 restatement of the reference
algorithms (KernelMatrices.jl 0.9.0, Julia).  It exists so that the tests, the
smoke check and the CPU-baseline leg of the benchmark have an independent
checker.  The product path (the HIP library and its Python/Julia bindings)
never imports it.

Modules
  kernels.py  covariance functions and their parameter derivatives
              (src/covariancefunctions.jl, README.md:42-48)
  hodlr.py    tree indices, ACA, Nystrom, KernelHODLR, symmetricfactorize!,
              factor apply, solve, logdet, matvec, derivative HODLR,
              stochastic gradient / Fisher / Hessian
              (src/factorizations.jl, src/utils.jl, src/HODLR/src/*.jl)
  dense.py    dense FP64 reference formulas (Cholesky logdet / solve, exact
              gradient) used to pin the restatement
  kdsort.py   KD-tree point ordering used to prepare sorted inputs

Parity pins (what checks each function; the tests live under tests/):
  kernels      closed forms vs mpmath besselk Matern (DLMF 10.39.2), finite
               differences for the parameter derivatives
  tree         brute-force enumeration of the dyadic split
  ACA          cross-approximation interpolation property, exact rank on
               rank-k matrices, two independent restatements (vectorised and
               scalar-loop) agree on ranks and pivots
  Nystrom      reproduces K(P, cols) at landmark rows; U V^T symmetric PSD
  factorize    W W^T == full(HK) (reference test/runtests.jl:28); level 0
               reduces to dense Cholesky
  solve/logdet dense FP64 Cholesky on full(HK) (test/runtests.jl:31-37)
  gradient     with vecs = sqrt(n) e_i the trace term is exact, so the
               stochastic gradient equals the dense formula
               0.5*(tr(K^-1 dK) - y' K^-1 dK K^-1 y) on full(HK)
"""
