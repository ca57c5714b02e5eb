// complex128 instantiation of the operator passes
#define TDDA_T double
#include "tdda_pass.inc"
