// mt_fused_q27_f64.hip -- the fused update kernels of D3Q27 with double storage.
#include "mt_fused.h"

namespace mt {
MT_FUSED_INSTANTIATE(27, double)
}  // namespace mt
