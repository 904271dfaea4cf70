// core_host.hpp — the per-width entry points of the host arithmetic (core_host_w.cpp) and their dispatch
// (core_host.cpp).  Operation codes follow limbforge::Operation (engine.hpp:13).
#pragma once

namespace lf {
enum : int {
    HOST_ADD = 0, HOST_SUB = 1, HOST_MUL = 2, HOST_DIV = 3, HOST_CADD = 4, HOST_CMUL = 5, HOST_CDIV = 6, HOST_SQUARE = 7,
    HOST_SQRT = 8, HOST_FMA = 9, HOST_FMS = 10, HOST_CFMA = 11, HOST_CFMS = 12, HOST_DOT2 = 100
};
// out = op(a, b, c) on public element records of N words; returns -1 for an unknown operation.
template <int N> int host_arithmetic(int op, const void* a, const void* b, const void* c, void* out);
// out = RN(x[0] x[1] + x[2] x[3] + x[4]), one rounding, on five records of N words
template <int N> void host_dot2(const void* const* x, void* out);
} // namespace lf
