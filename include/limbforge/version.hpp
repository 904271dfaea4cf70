#pragma once
// Library version of the MI355X build of the LimbForge/WolfNum API (the reference's
// include/limbforge/version.hpp contract: literal version macros, a header/binary check).
#define LIMBFORGE_VERSION_MAJOR 1
#define LIMBFORGE_VERSION_MINOR 3
#define LIMBFORGE_VERSION_PATCH 1
#define LIMBFORGE_API_VERSION LIMBFORGE_VERSION_MAJOR
#define LIMBFORGE_DETAIL_STRINGIFY_IMPL(x) #x
#define LIMBFORGE_DETAIL_STRINGIFY(x) LIMBFORGE_DETAIL_STRINGIFY_IMPL(x)
#define LIMBFORGE_VERSION_STRING                                                                                        \
    LIMBFORGE_DETAIL_STRINGIFY(LIMBFORGE_VERSION_MAJOR)                                                                 \
    "." LIMBFORGE_DETAIL_STRINGIFY(LIMBFORGE_VERSION_MINOR) "." LIMBFORGE_DETAIL_STRINGIFY(LIMBFORGE_VERSION_PATCH)
#define LIMBFORGE_VERSION_AT_LEAST(major, minor, patch)                                                                 \
    (LIMBFORGE_VERSION_MAJOR > (major) ||                                                                               \
     (LIMBFORGE_VERSION_MAJOR == (major) &&                                                                             \
      (LIMBFORGE_VERSION_MINOR > (minor) || (LIMBFORGE_VERSION_MINOR == (minor) && LIMBFORGE_VERSION_PATCH >= (patch)))))

namespace limbforge {
struct Version {
    unsigned major, minor, patch;
};
inline constexpr Version header_version{LIMBFORGE_VERSION_MAJOR, LIMBFORGE_VERSION_MINOR, LIMBFORGE_VERSION_PATCH};
inline constexpr unsigned api_version = LIMBFORGE_API_VERSION;
constexpr bool version_at_least(Version have, Version need) noexcept
{
    if (have.major != need.major) return have.major > need.major;
    if (have.minor != need.minor) return have.minor > need.minor;
    return have.patch >= need.patch;
}
// Same major and not older: source compatible.
constexpr bool version_compatible(Version have, Version need) noexcept
{
    return have.major == need.major && version_at_least(have, need);
}
Version library_version() noexcept;
const char* library_version_string() noexcept;
// Throws std::runtime_error on a major mismatch, an older library, or (exact) any difference.
void require_library_version(Version required, bool exact = false);
// Compares the caller's header version with the linked library.
inline void check_library_version()
{
    require_library_version({LIMBFORGE_VERSION_MAJOR, LIMBFORGE_VERSION_MINOR, LIMBFORGE_VERSION_PATCH}, true);
}
} // namespace limbforge
