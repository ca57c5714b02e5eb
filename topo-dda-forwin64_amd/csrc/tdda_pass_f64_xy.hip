// complex128 operator passes: y and x passes (P2, P3, P4)
#define TDDA_T double
#define TDDA_PASS_PART 2
#define TDDA_PXR_MID 1
#include "tdda_pass.inc"
