// Kernels for the 128-bit field.
#define SC_FIELD FieldP128
#define SC_LAUNCHERS launchers_p128
#include "inst_common.inc"
