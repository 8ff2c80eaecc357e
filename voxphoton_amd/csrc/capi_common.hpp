// Shared by the C ABI translation units: every entry point runs its body
// between VP_API_BEGIN and VP_API_END, which turn C++ exceptions into the
// status codes of voxphoton.h and record the message for vp_last_error().
#pragma once
#include <new>
#include <stdexcept>
#include <string>
#include "voxphoton.h"
#include "vp_internal.hpp"

namespace vp {
void set_error(const std::string& msg);
}

#define VP_API_BEGIN try {
#define VP_API_END                       \
  }                                      \
  catch (const vp::VpError& e) {         \
    vp::set_error(e.msg);                \
    return e.code;                       \
  }                                      \
  catch (const std::bad_alloc&) {        \
    vp::set_error("host allocation failed"); \
    return VP_ERR_OOM;                   \
  }                                      \
  catch (const std::exception& e) {      \
    vp::set_error(e.what());             \
    return VP_ERR_DEVICE;                \
  }                                      \
  catch (...) {                          \
    vp::set_error("unknown error");      \
    return VP_ERR_DEVICE;                \
  }                                      \
  vp::set_error("");                     \
  return VP_OK;
