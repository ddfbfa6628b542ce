// Field kernels for in_dim=2, coefficient GRID2, FP32 tables.
#define FF_INST_D 2
#define FF_INST_CK CK_GRID2
#define FF_INST_BF false
#include "ff_field_inst.inc"
