// Kernels for the 256-bit field, part 2 of 3 (single-workgroup tail kernels).
#define SC_FIELD FieldP256
#define SC_LAUNCHERS launchers_p256
#define SC_PART 2
#include "inst_common.inc"
