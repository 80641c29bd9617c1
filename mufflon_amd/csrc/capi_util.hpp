// Helpers of the extern "C" boundary shared by its translation units.
#pragma once
#include "../../include/mufflon_hip.h"

#include <string>

#include "mfh_internal.hpp"

struct mfh_context {
    mfh::Context c;
};

namespace mfh {

std::string& last_error();

// Run f, turning exceptions into a status code and the thread-local message
// (Mufflon's TRY / CATCH_ALL convention, interface.cpp:54-83).
template <class F>
int guarded(F&& f) {
    try {
        f();
        last_error().clear();
        return MFH_OK;
    } catch (const Error& e) {
        last_error() = e.what();
        return e.code;
    } catch (const std::exception& e) {
        last_error() = e.what();
        return MFH_ERROR_STATE;
    } catch (...) {
        last_error() = "unknown error";
        return MFH_ERROR_STATE;
    }
}

inline Context& ctx_of(mfh_context* ctx) {
    MFH_REQUIRE(ctx != nullptr, "null context");
    MFH_HIP(hipSetDevice(ctx->c.device));
    return ctx->c;
}

inline Lod& lod_of(Context& c, int32_t id) {
    MFH_REQUIRE(id >= 0 && id < (int32_t)c.lods.size() && c.lods[id], "unknown LoD id " + std::to_string(id));
    return *c.lods[id];
}

inline Scene& scene_of(Context& c, int32_t id) {
    MFH_REQUIRE(id >= 0 && id < (int32_t)c.scenes.size() && c.scenes[id], "unknown scene id " + std::to_string(id));
    return *c.scenes[id];
}

}  // namespace mfh
