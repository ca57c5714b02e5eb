// complex64 operator passes: fused z + y passes (P1 + P2, P4 + P5)
#define TDDA_T float
#define TDDA_PASS_PART 3
#include "tdda_pass.inc"
