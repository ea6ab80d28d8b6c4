// Fused stepping kernel instantiations: PrecF32, Burgers.
#include "pml_fdm_kernels.h"

namespace pml {
PML_DEFINE_TABLE_E(fdm_table_f32_e, PrecF32)
}  // namespace pml
