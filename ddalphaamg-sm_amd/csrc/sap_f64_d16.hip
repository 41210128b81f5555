// Coarse SAP kernels for double, dof = 16 (see sap_impl.h).
#define DD_SAPC_UNIT
#include "sap_impl.h"

namespace dd {
template void dispatch_sapc<double, 16>(int, hipStream_t, const _Float16*, const cplx<double>*, const cplx<double>*, cplx<double>*, int, int,
                                      int, int, int, double, int, double, int, uint64_t, size_t, const int*);
template bool dispatch_sap_fused<double, 16>(hipStream_t, const cplx<double>*, const _Float16*, const cplx<double>*, cplx<double>*,
                                        int, int, int, int, double, int, double, int, uint64_t, size_t, int,
                                        const SapFuse&, int);
}  // namespace dd
