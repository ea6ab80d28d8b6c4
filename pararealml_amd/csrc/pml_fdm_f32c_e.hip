// Fused stepping kernel instantiations: PrecF32C, Burgers.
#include "pml_fdm_kernels.h"

namespace pml {
PML_DEFINE_TABLE_E(fdm_table_f32c_e, PrecF32C)
}  // namespace pml
