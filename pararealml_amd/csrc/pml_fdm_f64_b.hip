// Fused stepping kernel instantiations: PrecF64, shallow water / wave.
#include "pml_fdm_kernels.h"

namespace pml {
PML_DEFINE_TABLE_B(fdm_table_f64_b, PrecF64)
}  // namespace pml
