// Error type and status codes shared by device, host and ABI code (no HIP).
#pragma once

#include <stdexcept>
#include <string>

namespace gfz {

// Error type carried up to the C ABI, which turns it into a status code.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

enum Status : int {
  kOk = 0,
  kErrInvalid = -1,
  kErrHip = -2,
  kErrOutOfMemory = -3,
  kErrOverflow = -4,
  kErrCorrupt = -5,
  kErrBusy = -6,
};

}  // namespace gfz
