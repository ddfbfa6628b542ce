// Field kernels for in_dim=3, coefficient GRID3, BF16 tables.
#define FF_INST_D 3
#define FF_INST_CK CK_GRID3
#define FF_INST_BF true
#include "ff_field_inst.inc"
