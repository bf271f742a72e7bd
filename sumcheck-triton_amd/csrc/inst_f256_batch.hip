// Kernels for the 256-bit field, part 3 of 3 (batched stages, wide compositions).
#define SC_FIELD FieldP256
#define SC_LAUNCHERS launchers_p256
#define SC_PART 3
#include "inst_common.inc"
