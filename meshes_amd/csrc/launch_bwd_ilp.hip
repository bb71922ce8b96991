// launch_bwd_ilp.hip -- the backdiff tiles of ilp_tile() (tiles.h), built with the iterative-ilp
// machine scheduler (Makefile).
#define MP_ILP_SET 1
#include "launch_bwd.hip"
