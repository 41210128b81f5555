// Wilson-Dirac operator on halo-padded host spinors (include/dirac_operator.h:15-31),
// applied on the GPU (dd_apply_D).
#ifndef DIRAC_OPERATOR_INCLUDED
#define DIRAC_OPERATOR_INCLUDED
#include "boundary.h"
#include "utils.h"

void D_phi(const spinor& U, const spinor& phi, spinor& Dphi, const double& m0);
void D_dagger_phi(const spinor& U, const spinor& phi, spinor& Dphi, const double& m0);
void D_D_dagger_phi(const spinor& U, const spinor& phi, spinor& Dphi, const double& m0);


#endif
