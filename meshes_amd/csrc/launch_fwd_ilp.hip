// launch_fwd_ilp.hip -- the fwdprop tiles of ilp_fwd_tile() (tiles.h), built with the iterative-ilp
// machine scheduler (Makefile).
#define MP_ILP_SET 1
#include "launch_fwd.hip"
