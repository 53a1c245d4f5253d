// hc_cell_f32.hip -- the fused per-cell kernels of the FP32-shifted form
// (hc_cell_body.h), compiled with multiply-add contraction.
#include "hc_cell_body.h"

namespace hcg {
template hipError_t launch_cell_step_multi<float>(hipStream_t, const TypeDev&, const CellCase*, int, int, int, int);
template hipError_t launch_cell_step<float>(hipStream_t, const VertexDev&, const TypeDev&, const int32_t*, int,
                                            const int64_t*, const float*, const uint8_t*, const LatticeDev&, float*,
                                            const double*, double*, double, int, int, int, unsigned long long*, int,
                                            bool);
template hipError_t launch_cell_step_mixed<float>(hipStream_t, const VertexDev&, const TypeDev*, const TypeDev*, int,
                                              const int*, const int2*, int, unsigned int*, int, const int64_t*,
                                              const float*, const uint8_t*, const LatticeDev&, float*, const double*,
                                              double*, double, int, int, unsigned long long*);
}  // namespace hcg
