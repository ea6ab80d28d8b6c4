// Fused stepping kernel instantiations: PrecF64, Burgers.
#include "pml_fdm_kernels.h"

namespace pml {
PML_DEFINE_TABLE_E(fdm_table_f64_e, PrecF64)
}  // namespace pml
