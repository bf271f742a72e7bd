// Kernels for the 256-bit field.
#define SC_FIELD FieldP256
#define SC_LAUNCHERS launchers_p256
#include "inst_common.inc"
