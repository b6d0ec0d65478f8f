// Test infrastructure: the host set-up built alone, with the sanitizers (see
// tests/test_sanitized_setup.py). The device passes (csrc/setup_device.hip) are not in that
// build; asking for them fails loudly.
#include <stdexcept>

#include "setup_device.h"

namespace asdev {

void set_up(const Input &, Output &) {
    throw std::runtime_error("the device set-up is not in the sanitized host build");
}

}  // namespace asdev
