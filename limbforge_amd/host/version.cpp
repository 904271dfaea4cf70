// Library version of the facade (the contract of the reference's src/version.cpp).
#include "limbforge/version.hpp"

#include <stdexcept>
#include <string>

namespace limbforge {
Version library_version() noexcept { return header_version; }
const char* library_version_string() noexcept { return LIMBFORGE_VERSION_STRING; }
void require_library_version(Version required, bool exact)
{
    const Version have = library_version();
    const bool same_release = have.minor == required.minor && have.patch == required.patch;
    if (version_compatible(have, required) && (!exact || same_release)) return;
    throw std::runtime_error(std::string("LimbForge library version ") + library_version_string() +
                             (exact ? " does not match headers " : " does not satisfy required API version ") +
                             std::to_string(required.major) + "." + std::to_string(required.minor) + "." + std::to_string(required.patch));
}
} // namespace limbforge
