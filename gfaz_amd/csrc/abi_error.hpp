// Error plumbing of the C ABI: nothing throws across it; every entry point
// catches, records the message for gfz_last_error() and returns a status code.
#pragma once

#include "error.hpp"
#include "gfaz_amd.h"

#include <new>
#include <string>

namespace gfz {

extern thread_local std::string g_last_error;  // defined in capi.hip

inline int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

template <class F>
int guarded(const char *prefix, F &&f) {
  try {
    f();
    return GFZ_OK;
  } catch (const Error &e) {
    return fail(e.code, std::string(prefix) + e.what());
  } catch (const std::bad_alloc &) {
    return fail(GFZ_E_NOMEM, std::string(prefix) + "host allocation failed");
  } catch (const std::exception &e) {
    return fail(GFZ_E_INVALID, std::string(prefix) + e.what());
  }
}

}  // namespace gfz
