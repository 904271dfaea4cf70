// C++ API parity test of the Engine facade on the GPU, written as the reference's own tests use the API
// (tests/test_buffers.cpp, test_broadcast.cpp, test_cast.cpp, examples/resident.cpp): host-array runs
// against the MPFR golden fixtures, the resident 16-step chain through CommandBatch::run, broadcast
// operands, casts against the oracle, mapped buffers, and the ownership / busy / submission rules.
// Built and run by tests/test_cpp_api.py (compiled on the CPU; run on the GPU).
#include <limbforge/engine.hpp>

#include "lf_oracle.h"

#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <random>
#include <string>
#include <vector>

using namespace limbforge;

static int failures = 0;
#define REQUIRE(cond, msg)                                                                                             \
    do {                                                                                                               \
        if (!(cond)) {                                                                                                 \
            ++failures;                                                                                                \
            std::cerr << "FAIL " << msg << " (" << __FILE__ << ":" << __LINE__ << ")\n";                               \
        }                                                                                                              \
    } while (0)

template <class E> static bool throws(const std::function<void()>& f)
{
    try {
        f();
    } catch (const E&) {
        return true;
    } catch (...) {
        return false;
    }
    return false;
}

struct Golden {
    std::vector<unsigned char> bytes;
    size_t off = 16;
    explicit Golden(const std::string& path)
    {
        std::ifstream f(path, std::ios::binary);
        bytes.assign(std::istreambuf_iterator<char>(f), {});
    }
    uint32_t u32(size_t at) const
    {
        uint32_t v;
        std::memcpy(&v, &bytes[at], 4);
        return v;
    }
    uint32_t header(int i) const { return u32(4 + 4 * i); }
    const unsigned char* take(size_t n)
    {
        const unsigned char* p = &bytes[off];
        off += n;
        return p;
    }
};

static std::string dir;

template <int Bits> void golden_elementwise(Engine& e)
{
    Golden g(dir + "/elementwise_" + std::to_string(Bits) + ".bin");
    REQUIRE(!g.bytes.empty(), "fixture " << Bits);
    const size_t E = Bits / 8 + 12;
    for (uint32_t s = 0; s < g.header(2); ++s) {
        const uint32_t op = g.u32(g.off), count = g.u32(g.off + 4), has_c = g.u32(g.off + 8);
        g.off += 16;
        const size_t rec = E * (op >= 4 && op <= 6 ? 2 : 1);
        const unsigned char* a = g.take(count * rec);
        const unsigned char* b = g.take(count * rec);
        const unsigned char* c = has_c ? g.take(count * rec) : nullptr;
        const unsigned char* want = g.take(count * rec);
        std::vector<unsigned char> out(count * rec);
        Operation o = Operation(op);
        if (operation_is_unary(o)) e.run_unary(Bits, o, a, out.data(), count);
        else if (operation_is_ternary(o)) e.run_ternary(Bits, o, a, b, c, out.data(), count);
        else e.run(Bits, o, a, b, out.data(), count);
        REQUIRE(std::memcmp(out.data(), want, out.size()) == 0, "Engine::run golden bits=" << Bits << " op=" << op);
    }
}

template <int Bits> void resident_chain(Engine& e, const unsigned char* x0, const unsigned char* f, const unsigned char* want, size_t count, int steps)
{
    using F = Float<Bits>;
    auto x = e.make_buffer<F>(count), y = e.make_buffer<F>(count);
    x.upload(reinterpret_cast<const F*>(x0), count);
    y.upload(reinterpret_cast<const F*>(f), count);
    auto batch = e.batch();
    for (int i = 0; i < steps; ++i) batch.run(Operation::mul, x, y, x);
    auto submission = batch.submit();
    Timing t = submission.wait();
    Timing t2 = submission.wait();
    REQUIRE(t.gpu_seconds == t2.gpu_seconds && t.wall_seconds == t2.wall_seconds, "repeated wait returns the same timing");
    std::vector<F> out(count);
    x.download(out.data(), count);
    REQUIRE(std::memcmp(out.data(), want, count * sizeof(F)) == 0, "resident chain bits=" << Bits);
}

void golden_chains(Engine& e)
{
    Golden g(dir + "/chain.bin");
    for (uint32_t s = 0; s < g.header(2); ++s) {
        const uint32_t bits = g.u32(g.off), count = g.u32(g.off + 4), steps = g.u32(g.off + 8);
        g.off += 16;
        const size_t E = bits / 8 + 12;
        const unsigned char* x = g.take(count * E);
        const unsigned char* f = g.take(count * E);
        const unsigned char* want = g.take(count * E);
        if (bits == 1024) resident_chain<1024>(e, x, f, want, count, steps);
        else if (bits == 256) resident_chain<256>(e, x, f, want, count, steps);
        else if (bits == 384) resident_chain<384>(e, x, f, want, count, steps);
    }
}

template <int Bits> Float<Bits> random_value(std::mt19937_64& rng, int span)
{
    Float<Bits> x = zero<Bits / 32>();
    for (auto& l : x.limb) l = uint32_t(rng());
    x.limb[Bits / 32 - 1] |= 0x80000000u;
    x.sign = (rng() & 1) ? 1 : -1;
    x.exponent = int(rng() % (2 * span + 1)) - span;
    return x;
}

void broadcast_and_mixed_batch(Engine& e)
{
    constexpr int Bits = 512;
    using F = Float<Bits>;
    std::mt19937_64 rng(7);
    const size_t n = 1000;
    std::vector<F> a(n), b(64), c(n), out(n), want(n), tmp(n);
    for (auto& v : a) v = random_value<Bits>(rng, 100);
    for (auto& v : b) v = random_value<Bits>(rng, 100);
    for (auto& v : c) v = random_value<Bits>(rng, 100);
    auto A = e.make_buffer<F>(n), B = e.make_buffer<F>(64), C = e.make_buffer<F>(n), O = e.make_buffer<F>(n);
    A.upload(a.data(), n);
    B.upload(b.data(), 64);
    C.upload(c.data(), n);
    auto batch = e.batch();
    batch.run(Operation::fma, A, B, C, O, Broadcast{1, 64}, Broadcast{1, 0});
    batch.run(Operation::div, O, B, O, Broadcast{16, 0});
    batch.run(Operation::sqrt, O, O);
    batch.submit().wait();
    O.download(out.data(), n);
    lfo_elementwise(Bits, 9, a.data(), b.data(), c.data(), tmp.data(), n, 1, 64, 1, 0);
    lfo_elementwise(Bits, 3, tmp.data(), b.data(), nullptr, want.data(), n, 16, 0, 1, 0);
    lfo_elementwise(Bits, 8, want.data(), nullptr, nullptr, tmp.data(), n, 1, 0, 1, 0);
    REQUIRE(std::memcmp(out.data(), tmp.data(), n * sizeof(F)) == 0, "fma/div/sqrt batch with broadcasts");
}

void casts(Engine& e)
{
    std::mt19937_64 rng(11);
    const size_t n = 777;
    std::vector<Float<1024>> wide(n);
    for (auto& v : wide) v = random_value<1024>(rng, 1000);
    wide[3] = zero<32>(invalid);
    wide[4].exponent = 1000000000;
    for (auto& l : wide[4].limb) l = 0xffffffffu; // narrowing carries past the top exponent
    std::vector<Float<96>> narrow(n), want(n);
    e.cast(1024, 96, false, wide.data(), narrow.data(), n);
    lfo_cast(1024, 96, 0, wide.data(), want.data(), n);
    REQUIRE(std::memcmp(narrow.data(), want.data(), n * sizeof(Float<96>)) == 0, "host cast 1024 -> 96");
    REQUIRE(narrow[4].status == exponent_overflow, "narrowing overflow status");
    // resident widen: exact
    auto N = e.make_buffer<Float<96>>(n);
    auto W = e.make_buffer<Float<640>>(n);
    N.upload(narrow.data(), n);
    auto batch = e.batch();
    batch.cast(N, W);
    batch.submit().wait();
    std::vector<Float<640>> w(n), ww(n);
    W.download(w.data(), n);
    lfo_cast(96, 640, 0, narrow.data(), ww.data(), n);
    REQUIRE(std::memcmp(w.data(), ww.data(), n * sizeof(Float<640>)) == 0, "resident cast 96 -> 640");
    // complex
    std::vector<Complex<8>> zc(n);
    std::vector<Complex<4>> zn(n), zw(n);
    for (auto& z : zc) z = {random_value<256>(rng, 50), random_value<256>(rng, 50)};
    e.cast(256, 128, true, zc.data(), zn.data(), n);
    lfo_cast(256, 128, 1, zc.data(), zw.data(), n);
    REQUIRE(std::memcmp(zn.data(), zw.data(), n * sizeof(Complex<4>)) == 0, "complex cast 256 -> 128");
}

void ownership(Engine& e)
{
    using F = Float<128>;
    Engine other;
    auto x = e.make_buffer<F>(8), y = e.make_buffer<F>(8), z = other.make_buffer<F>(8);
    std::vector<F> v(8);
    std::mt19937_64 rng(3);
    for (auto& a : v) a = random_value<128>(rng, 10);
    x.upload(v.data(), 8);
    y.upload(v.data(), 8);
    {
        auto batch = e.batch();
        REQUIRE(throws<std::invalid_argument>([&] { batch.run(Operation::mul, x, z, x); }), "foreign buffer rejected");
    }
    auto batch = e.batch();
    batch.run(Operation::add, x, y, x);
    auto sub = batch.submit();
    REQUIRE(throws<std::logic_error>([&] { batch.submit(); }), "second submit rejected");
    REQUIRE(throws<std::logic_error>([&] { batch.run(Operation::add, x, y, x); }), "encode after submit rejected");
    sub.wait();
    REQUIRE(sub.ready(), "waited submission is ready");
    // mapped(): host writes reach the next batch
    F* m = y.mapped();
    m[0] = zero<4>();
    m[0].limb[3] = 0x80000000u;
    m[0].sign = 1; // y[0] = 1
    auto b2 = e.batch();
    b2.run(Operation::mul, x, y, x);
    b2.submit().wait();
    std::vector<F> got(8);
    x.download(got.data(), 8);
    std::vector<F> sum(8);
    lfo_elementwise(128, 0, v.data(), v.data(), nullptr, sum.data(), 8, 1, 0, 1, 0);
    REQUIRE(std::memcmp(&got[0], &sum[0], sizeof(F)) == 0, "mapped write-back used by the next batch (x*1 = x)");
    REQUIRE(throws<std::invalid_argument>([&] { x.upload(v.data(), 9); }), "upload beyond size rejected");
    Submission empty;
    REQUIRE(throws<std::logic_error>([&] { empty.wait(); }), "empty submission");
    REQUIRE(throws<std::invalid_argument>([&] { e.run(80, Operation::mul, v.data(), v.data(), got.data(), 8); }), "bad width");
    REQUIRE(throws<std::invalid_argument>([&] { e.run(128, Operation::mul, v.data(), v.data(), got.data(), 8, Broadcast{0, 0}); }),
            "zero stride");
}

int main(int argc, char** argv)
{
    dir = argc > 1 ? argv[1] : "tests/golden";
    try {
        check_library_version();
        Engine engine;
        std::cout << "device: " << engine.device_name() << "\n";
        golden_elementwise<64>(engine);
        golden_elementwise<96>(engine);
        golden_elementwise<256>(engine);
        golden_elementwise<384>(engine);
        golden_elementwise<512>(engine);
        golden_elementwise<1024>(engine);
        golden_chains(engine);
        broadcast_and_mixed_batch(engine);
        casts(engine);
        ownership(engine);
    } catch (const std::exception& ex) {
        std::cerr << "exception: " << ex.what() << "\n";
        return 2;
    }
    std::cout << (failures ? "FAILED " : "passed ") << failures << " failures\n";
    return failures ? 1 : 0;
}
