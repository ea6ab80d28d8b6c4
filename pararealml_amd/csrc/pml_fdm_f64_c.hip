// Fused stepping kernel instantiations: PrecF64, Navier-Stokes.
#include "pml_fdm_kernels.h"

namespace pml {
PML_DEFINE_TABLE_C(fdm_table_f64_c, PrecF64)
}  // namespace pml
