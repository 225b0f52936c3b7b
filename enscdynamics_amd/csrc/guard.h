// Exceptions stop at the C ABI: a host allocation failure or a standard
// library exception inside an entry point becomes an ensc_status and a
// message for ensc_last_error(), instead of unwinding into C callers (or
// ending the process from a worker thread).
#pragma once

#include <exception>
#include <new>
#include <string>

#include "../../include/ensc_engine.h"

namespace ensc {
extern thread_local std::string g_last_error;

template <class F>
ensc_status guarded(F&& f) noexcept {
    try {
        return f();
    } catch (const std::bad_alloc&) {
        g_last_error = "host memory exhausted";
        return ENSC_ERR_MEM;
    } catch (const std::exception& ex) {
        g_last_error = std::string("unexpected host error: ") + ex.what();
        return ENSC_ERR_STATE;
    } catch (...) {
        g_last_error = "unexpected host error";
        return ENSC_ERR_STATE;
    }
}
}  // namespace ensc
