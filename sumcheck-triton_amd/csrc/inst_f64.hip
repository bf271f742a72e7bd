// Kernels for the 64-bit field.
#define SC_FIELD FieldP64
#define SC_LAUNCHERS launchers_p64
#include "inst_common.inc"
