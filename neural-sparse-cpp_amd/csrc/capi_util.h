// Exception -> status-code translation for the C ABI.
#pragma once

#include <exception>
#include <new>
#include <stdexcept>
#include <string>

#include "nsparse_c.h"

namespace nsparse_amd {

// A failing HIP runtime call. Kept apart from std::runtime_error so callers can
// tell a device fault from a corrupt file.
struct device_error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void set_last_error(const char* msg);

template <class F>
int guarded(F&& f) noexcept {
    try {
        f();
        set_last_error("");
        return NSPARSE_OK;
    } catch (const std::invalid_argument& e) {
        set_last_error(e.what());
        return NSPARSE_ERR_INVALID_ARGUMENT;
    } catch (const std::out_of_range& e) {
        set_last_error(e.what());
        return NSPARSE_ERR_OUT_OF_RANGE;
    } catch (const device_error& e) {
        set_last_error(e.what());
        return NSPARSE_ERR_DEVICE;
    } catch (const std::bad_alloc& e) {
        set_last_error(e.what());
        return NSPARSE_ERR_BAD_ALLOC;
    } catch (const std::runtime_error& e) {
        set_last_error(e.what());
        return NSPARSE_ERR_RUNTIME;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return NSPARSE_ERR_UNKNOWN;
    } catch (...) {
        set_last_error("unknown exception");
        return NSPARSE_ERR_UNKNOWN;
    }
}

}  // namespace nsparse_amd
