// Kernels for the 256-bit field, part 1 of 3 (pass kernels, small kernels, launcher table).
#define SC_FIELD FieldP256
#define SC_LAUNCHERS launchers_p256
#define SC_PART 1
#include "inst_common.inc"
