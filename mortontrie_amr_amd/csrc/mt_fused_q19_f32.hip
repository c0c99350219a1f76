// mt_fused_q19_f32.hip -- the fused update kernels of D3Q19 with float storage.
#include "mt_fused.h"

namespace mt {
MT_FUSED_INSTANTIATE(19, float)
}  // namespace mt
