// Krylov comparison solvers of the reference (include/conjugate_gradient.h:14-16), run on
// the GPU (dd_solve_cg, dd_solve_bicgstab).  Knobs: CG::max_iter/tol, BiCG::max_iter/tol.
#ifndef CONJUGATE_GRADIENT_H
#define CONJUGATE_GRADIENT_H
#include "dirac_operator.h"

// D^dagger D x = phi from x = phi; returns the iteration count, 0 if not converged
// (src/conjugate_gradient.cpp:4-84)
int conjugate_gradient(const spinor& U, const spinor& phi, spinor& x, const double& m0, const bool print);
// D x = phi from x0; returns the iteration count, BiCG::max_iter if not converged
// (src/conjugate_gradient.cpp:87-176)
int bi_cgstab(const spinor& U, const spinor& phi, const spinor& x0, spinor& x, const double& m0,
              const bool& print_message);

#endif
