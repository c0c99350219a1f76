// mt_fused_q19_f16.hip -- the fused update kernels of D3Q19 with _Float16 storage.
#include "mt_fused.h"

namespace mt {
MT_FUSED_INSTANTIATE(19, _Float16)
}  // namespace mt
