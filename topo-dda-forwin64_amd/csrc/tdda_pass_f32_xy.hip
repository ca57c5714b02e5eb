// complex64 operator passes: y and x passes (P2, P3, P4)
#define TDDA_T float
#define TDDA_PASS_PART 2
#include "tdda_pass.inc"
