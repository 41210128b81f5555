// SAP block-solve kernels instantiated for float (see sap_impl.h).
#include "sap_impl.h"

namespace dd {
template void sap_blocks0<float>(hipStream_t, const cplx<float>*, const cplx<float>*, cplx<float>*, int, int, int, int, int,
                               double, int, int, double, int, uint64_t, size_t, const int*, bool);
template void sap_blocksc<float>(hipStream_t, const cplx<float>*, const cplx<float>*, cplx<float>*, int, int, int, int, int,
                               int, double, int, int, double, int, uint64_t, size_t, const int*, int,
                                 const _Float16*);
template std::vector<int> sapc_variants<float>();
template bool sapc_supported<float>(int, int, int);
template bool sap_fused_iteration<float>(hipStream_t, const cplx<float>*, const _Float16*, const cplx<float>*, cplx<float>*,
                                      int, int, int, int, int, double, int, int, double, int, uint64_t, size_t, int,
                                      const SapFuse&, int);
}  // namespace dd
