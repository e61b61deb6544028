// bodies.hip -- flow-move, vortex shedding, the boundary-condition SLAE and
// the force integrals (entry points; kernels follow).
#include "api.hpp"

using namespace vvhd;

extern "C" {

int vvhd_move_and_clean(vvhd_ctx* ctx, int remove, int64_t* cleaned) {
    REQUIRE_CTX(ctx);
    (void)remove;
    (void)cleaned;
    ctx->c.err = "move_and_clean: not available in this build";
    return VVHD_EUNSUPPORTED;
}

int vvhd_vortex_shed(vvhd_ctx* ctx) {
    REQUIRE_CTX(ctx);
    ctx->c.err = "vortex_shed: not available in this build";
    return VVHD_EUNSUPPORTED;
}

int vvhd_calc_circulation(vvhd_ctx* ctx, double inf_x, double inf_y) {
    REQUIRE_CTX(ctx);
    (void)inf_x;
    (void)inf_y;
    ctx->c.err = "calc_circulation: not available in this build";
    return VVHD_EUNSUPPORTED;
}

int vvhd_calc_forces(vvhd_ctx* ctx) {
    REQUIRE_CTX(ctx);
    ctx->c.err = "calc_forces: not available in this build";
    return VVHD_EUNSUPPORTED;
}

int vvhd_zero_forces(vvhd_ctx* ctx) {
    REQUIRE_CTX(ctx);
    ctx->c.err = "zero_forces: not available in this build";
    return VVHD_EUNSUPPORTED;
}

}  // extern "C"
