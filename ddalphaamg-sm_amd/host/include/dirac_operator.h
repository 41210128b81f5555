// Wilson-Dirac operator on halo-padded host spinors (include/dirac_operator.h:15-31),
// applied on the GPU (dd_apply_D).
#ifndef DIRAC_OPERATOR_INCLUDED
#define DIRAC_OPERATOR_INCLUDED
#include "variables.h"

void D_phi(const spinor& U, const spinor& phi, spinor& Dphi, const double& m0);
void D_dagger_phi(const spinor& U, const spinor& phi, spinor& Dphi, const double& m0);
void D_D_dagger_phi(const spinor& U, const spinor& phi, spinor& Dphi, const double& m0);

// halo-padded site index of interior (x, t), x in [1, Nx], t in [1, Nt] (include/utils.h:46-52)
inline int idx(int x, int t, int mu) { return ((x * (mpi::width_t + 2) + t) * LV::dof + mu); }

#endif
