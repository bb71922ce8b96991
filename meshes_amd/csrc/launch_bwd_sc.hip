// launch_bwd_sc.hip -- the scaled-unitary backdiff kernels (generic MZI crossings asserted to
// be c U, MESHPROP_SCALED_UNITARY), every complex tile.
#define MP_SC_SET 1
#include "launch_bwd.hip"
