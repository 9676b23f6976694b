// lbib_mixed.hip -- mixed-precision lattice kernels (FP32 arithmetic, 16-bit
// deviation storage).  Placeholder launchers until the kernels land.
#include "lbib_internal.h"

namespace lbib {
int launch_init_mixed(lbib_ctx* c, const double*) {
    c->err = "mixed-precision storage is not available in this build";
    return -1;
}
int launch_stream_collide_mixed(lbib_ctx* c, uint32_t, bool) {
    c->err = "mixed-precision storage is not available in this build";
    return -1;
}
int launch_pack(lbib_ctx* c, const double*, int, int) {
    c->err = "mixed-precision storage is not available in this build";
    return -1;
}
int launch_unpack(lbib_ctx* c, double*, int, int) {
    c->err = "mixed-precision storage is not available in this build";
    return -1;
}
}  // namespace lbib
