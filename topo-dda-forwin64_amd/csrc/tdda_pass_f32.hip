// complex64 instantiation of the operator passes
#define TDDA_T float
#include "tdda_pass.inc"
