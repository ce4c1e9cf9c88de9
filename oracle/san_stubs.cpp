// san_stubs.cpp — for the sanitizer build of the OBJ parser only
// (oracle/Makefile `sanitize`): rp_obj.cpp reports errors through the
// library's error slot, which lives in rp_world.cpp (HIP host code, not part
// of this CPU build). Test infrastructure.
#include <cstdint>
#include <string>

namespace rp {
void rp_set_last_error(uint32_t, const std::string&) {}
}  // namespace rp
