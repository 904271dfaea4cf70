// Consumer of the installed LimbForge package: checks the header/library version, then runs the host scalar
// arithmetic (limbforge::add ... cdiv) and the host exact dot (exact_dot_add) on records read from argv[1]
// and writes the results to argv[2]; tests/test_package_consumer.py compares them with the oracle.
// Input (uint32 words): C; then per width N in {2, 3, 8, 16, 32}: C records each of a, b, c; then D and per
// width D dot problems {K, subtract, has_c, c, a[K], b[K]}.  Output per width: for each case
// add, sub, mul, div, sqrt(a), fma, fms, square(a), then cadd, cmul, cdiv of x = (a, b) and y = (c, b), cfma and
// cfms of x, y and z = (c, a); then
// the D dot results.
#include <limbforge/core.hpp>
#include <limbforge/linalg.hpp>
#include <limbforge/version.hpp>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

using namespace limbforge;

static std::vector<std::uint32_t> in;
static std::size_t pos = 0;
static std::vector<std::uint32_t> out;

template <int N> Number<N> take()
{
    Number<N> r;
    if (pos + sizeof r / 4 > in.size()) throw std::out_of_range("input file too short");
    std::memcpy(&r, in.data() + pos, sizeof r);
    pos += sizeof r / 4;
    return r;
}
template <class T> void give(const T& r)
{
    const std::uint32_t* w = reinterpret_cast<const std::uint32_t*>(&r);
    out.insert(out.end(), w, w + sizeof(T) / 4);
}

template <int N> void width(std::uint32_t C)
{
    std::vector<Number<N>> a(C), b(C), c(C);
    for (auto& x : a) x = take<N>();
    for (auto& x : b) x = take<N>();
    for (auto& x : c) x = take<N>();
    for (std::uint32_t i = 0; i < C; ++i) {
        give(add(a[i], b[i]));
        give(sub(a[i], b[i]));
        give(mul(a[i], b[i]));
        give(div(a[i], b[i]));
        give(limbforge::sqrt(a[i]));
        give(limbforge::fma(a[i], b[i], c[i]));
        give(fms(a[i], b[i], c[i]));
        give(square(a[i]));
        const Complex<N> x{a[i], b[i]}, y{c[i], b[i]};
        give(cadd(x, y));
        give(cmul(x, y));
        give(cdiv(x, y));
        const Complex<N> z{c[i], a[i]};
        give(cfma(x, y, z));
        give(cfms(x, y, z));
    }
}

template <int N> void dots(std::uint32_t D)
{
    for (std::uint32_t d = 0; d < D; ++d) {
        const std::uint32_t K = in.at(pos), subtract = in.at(pos + 1), has_c = in.at(pos + 2);
        pos += 3;
        const Number<N> c = take<N>();
        std::vector<Number<N>> a(K), b(K);
        for (auto& x : a) x = take<N>();
        for (auto& x : b) x = take<N>();
        give(exact_dot_add<N>(has_c ? &c : nullptr, subtract != 0, a.data(), 1, b.data(), 1, K));
    }
}

int main(int argc, char** argv)
{
    if (std::string(library_version_string()) != LF_EXPECTED_VERSION) return 1;
    check_library_version();
    if (argc < 3) return 2;
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 3;
    std::uint32_t w;
    while (std::fread(&w, 4, 1, f) == 1) in.push_back(w);
    std::fclose(f);
    const std::uint32_t C = in.at(pos++);
    width<2>(C);
    width<3>(C);
    width<8>(C);
    width<16>(C);
    width<32>(C);
    const std::uint32_t D = in.at(pos++);
    dots<2>(D);
    dots<3>(D);
    dots<8>(D);
    dots<16>(D);
    dots<32>(D);
    FILE* g = std::fopen(argv[2], "wb");
    if (!g || std::fwrite(out.data(), 4, out.size(), g) != out.size()) return 4;
    std::fclose(g);
    std::printf("package consumer %s: %zu result words\n", library_version_string(), out.size());
    return 0;
}
