// Library-level C ABI: version, errors, device selection.
#include "common.hpp"

namespace rnng {
const char* last_error();
int init(int device);
}  // namespace rnng

extern "C" {

int rnng_version(void) { return 100; }

const char* rnng_last_error(void) { return rnng::last_error(); }

int rnng_init(int device) { return rnng::init(device); }

}  // extern "C"
