// Coarse SAP kernels for float, dof = 32 (see sap_impl.h).
#define DD_SAPC_UNIT
#include "sap_impl.h"

namespace dd {
template void dispatch_sapc<float, 32>(int, hipStream_t, const _Float16*, const cplx<float>*, const cplx<float>*, cplx<float>*, int, int,
                                      int, int, int, double, int, double, int, uint64_t, size_t, const int*);
template bool dispatch_sap_fused<float, 32>(hipStream_t, const cplx<float>*, const _Float16*, const cplx<float>*, cplx<float>*,
                                        int, int, int, int, double, int, double, int, uint64_t, size_t, int,
                                        const SapFuse&, int);
}  // namespace dd
