// mt_fused_q19_f64.hip -- the fused update kernels of D3Q19 with double storage.
#include "mt_fused.h"

namespace mt {
MT_FUSED_INSTANTIATE(19, double)
}  // namespace mt
