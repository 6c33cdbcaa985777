// Instance shapes with compile-time instantiations of the cooperative kernel
// (name, nc, mc, mcp, ko, N, d, dense); any other shape runs on the
// run-time-shape kernel.  Each shape is built in its own translation unit
// pdq_coop_<name>.hip.
//   iso5    isotropic, nu = 4, d = 2: the batched Lotka-Volterra solves
//   iso4t6  isotropic, nu = 3, d = 2 with 6 tangent columns: LML gradients
//   dense5  dense, nu = 4, d = 2: single Lotka-Volterra solves
//   dense6  dense, nu = 5, d = 2: van der Pol with ts1
#pragma once
#define PDQ_FIXED_SHAPES(X)          \
  X(iso5, 5, 2, 2, 1, 5, 2, 0)       \
  X(iso4t6, 4, 14, 2, 1, 4, 2, 0)    \
  X(dense5, 10, 1, 1, 2, 5, 2, 1)    \
  X(dense6, 12, 1, 1, 2, 6, 2, 1)
