// Errors raised inside the library and turned into statuses at the C boundary.
// No device code: the job server's host core includes it on its own.
#pragma once

#include <stdexcept>
#include <string>

#include "menos_hip.h"

namespace mh {

struct Error : std::runtime_error {
    mh_status status;
    Error(mh_status s, const std::string& what) : std::runtime_error(what), status(s) {}
};

[[noreturn]] inline void fail(mh_status status, const std::string& what) { throw Error(status, what); }

}  // namespace mh
