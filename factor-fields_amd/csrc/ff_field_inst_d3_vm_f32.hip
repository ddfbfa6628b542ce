// Field kernels for in_dim=3, coefficient VM, FP32 tables.
#define FF_INST_D 3
#define FF_INST_CK CK_VM
#define FF_INST_BF false
#include "ff_field_inst.inc"
