// Fused stepping kernel instantiations: PrecF32C, Navier-Stokes.
#include "pml_fdm_kernels.h"

namespace pml {
PML_DEFINE_TABLE_C(fdm_table_f32c_c, PrecF32C)
}  // namespace pml
