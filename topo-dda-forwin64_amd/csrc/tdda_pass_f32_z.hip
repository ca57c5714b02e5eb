// complex64 operator passes: z passes (P1, P5)
#define TDDA_T float
#define TDDA_PASS_PART 1
#include "tdda_pass.inc"
