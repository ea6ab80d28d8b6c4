// Fused stepping kernel instantiations: PrecF32C, shallow water / wave.
#include "pml_fdm_kernels.h"

namespace pml {
PML_DEFINE_TABLE_B(fdm_table_f32c_b, PrecF32C)
}  // namespace pml
