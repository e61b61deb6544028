// api.hpp -- the opaque context behind the C ABI and the exception-to-status
// conversion shared by the translation units that implement entry points.
#pragma once

#include "common.hpp"

struct vvhd_ctx {
    vvhd::Ctx c;
};

#define API_BEGIN try {
#define API_END(ctx)                  \
    }                                 \
    catch (const vvhd::Error& e) {    \
        (ctx)->c.err = e.what();      \
        return e.code;                \
    }                                 \
    catch (const std::exception& e) { \
        (ctx)->c.err = e.what();      \
        return VVHD_EHIP;             \
    }                                 \
    return VVHD_OK;

#define REQUIRE_CTX(ctx)      \
    if (!(ctx)) return VVHD_EINVAL; \
    ++(ctx)->c.api_calls

// Entry points whose reference counterpart couples the whole Space refuse to
// run while several simulations share the context.
#define REQUIRE_ONE_SIM(c, name) \
    VVHD_CHECK((c).nsim == 1, VVHD_EUNSUPPORTED, name ": not available while several simulations share the context")
