// complex64 operator passes: y and x passes (P2, P3, P4), with the
// register-resident kernels also for x / y lengths above 256
#define TDDA_T float
#define TDDA_PASS_PART 2
#define TDDA_PXR_LONG 1
#include "tdda_pass.inc"
