// Fused stepping kernel instantiations: PrecF32, Navier-Stokes.
#include "pml_fdm_kernels.h"

namespace pml {
PML_DEFINE_TABLE_C(fdm_table_f32_c, PrecF32)
}  // namespace pml
