// Kernels for the 32-bit field.
#define SC_FIELD FieldP32
#define SC_LAUNCHERS launchers_p32
#include "inst_common.inc"
