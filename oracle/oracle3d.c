/* oracle3d.c — 3D oracle (restatement in progress). */
#include "oracle.h"
size_t orc3d_params_size(void) { return sizeof(orc_params3d); }
