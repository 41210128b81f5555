// SAP block-solve kernels instantiated for double (see sap_impl.h).
#include "sap_impl.h"

namespace dd {
template void sap_blocks0<double>(hipStream_t, const cplx<double>*, const cplx<double>*, cplx<double>*, int, int, int, int, int,
                               double, int, int, double, int, uint64_t, size_t, const int*, bool);
template void sap_blocksc<double>(hipStream_t, const cplx<double>*, const cplx<double>*, cplx<double>*, int, int, int, int, int,
                               int, double, int, int, double, int, uint64_t, size_t, const int*, int,
                                 const _Float16*);
template std::vector<int> sapc_variants<double>();
template bool sapc_supported<double>(int, int, int);
template bool sap_fused_iteration<double>(hipStream_t, const cplx<double>*, const _Float16*, const cplx<double>*, cplx<double>*,
                                      int, int, int, int, int, double, int, int, double, int, uint64_t, size_t, int,
                                      const SapFuse&, int);
}  // namespace dd
