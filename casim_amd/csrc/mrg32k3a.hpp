// mrg32k3a.hpp -- MRG32k3a in 32/64-bit integer arithmetic, host and device.
//
// The generator of CaSim's vendored RngStreams (rngstream.cpp).  The
// reference evaluates the recurrences in double precision; every
// intermediate is an exact integer there, so the integer form below yields
// the same states, and U01 / U01d are formed from them with the same double
// operations, which makes every draw bit-identical.
//
//   component 1: x_n = (1403580 x_{n-2} - 810728 x_{n-3}) mod m1, m1 = 2^32 - 209
//   component 2: y_n = (527612 y_{n-1} - 1370589 y_{n-3}) mod m2, m2 = 2^32 - 22853
//   U01  = ((x_n - y_n) mod m1 (as x > y ? x - y : x - y + m1)) / (m1 + 1)
//   draw = U01d: u = U01 + U01 * 2^-24, minus 1 if >= 1   (rngstream.cpp:271-283)
//
// State layout: s[0..2] = (x_{n-3}, x_{n-2}, x_{n-1}), s[3..5] likewise for y
// (RngStream::Cg).  A jump of k steps is s <- A^k s (mod m) per component.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define CASIM_HD __host__ __device__ __forceinline__
#else
#define CASIM_HD inline
#endif

namespace casim {
namespace mrg {

constexpr uint32_t M1 = 4294967087u;
constexpr uint32_t M2 = 4294944443u;
constexpr uint32_t C1 = 209u;    // 2^32 - M1
constexpr uint32_t C2 = 22853u;  // 2^32 - M2
constexpr uint32_t A12 = 1403580u;
constexpr uint32_t A13N = 810728u;
constexpr uint32_t A21 = 527612u;
constexpr uint32_t A23N = 1370589u;
constexpr double NORM = 1.0 / (4294967087.0 + 1.0);
constexpr double FACT = 5.9604644775390625e-8;  // 2^-24

// v mod (2^32 - c) for any 64-bit v and c < 2^15: fold the high word three
// times (2^32 = c mod m), then one conditional subtraction.
CASIM_HD uint32_t reduce(uint64_t v, uint32_t c, uint32_t m) {
    v = (v >> 32) * c + (v & 0xffffffffull);
    v = (v >> 32) * c + (v & 0xffffffffull);
    v = (v >> 32) * c + (v & 0xffffffffull);
    return (uint32_t)(v >= m ? v - m : v);
}

CASIM_HD uint32_t mulmod(uint32_t a, uint32_t b, uint32_t c, uint32_t m) {
    return reduce((uint64_t)a * b, c, m);
}

// One step of each component; returns the U01 value of the step.
CASIM_HD double step(uint32_t s[6]) {
    uint32_t p1 = reduce((uint64_t)mulmod(A12, s[1], C1, M1) + mulmod(M1 - A13N, s[0], C1, M1), C1, M1);
    s[0] = s[1];
    s[1] = s[2];
    s[2] = p1;
    uint32_t p2 = reduce((uint64_t)mulmod(A21, s[5], C2, M2) + mulmod(M2 - A23N, s[3], C2, M2), C2, M2);
    s[3] = s[4];
    s[4] = s[5];
    s[5] = p2;
    double d = (p1 > p2) ? ((double)p1 - (double)p2) : ((double)p1 - (double)p2 + 4294967087.0);
    return d * NORM;
}

// RandU01 of the vendored RngStream (incPrec = true): two steps per draw.
CASIM_HD double draw(uint32_t s[6]) {
    double u = step(s);
    u += step(s) * FACT;
    return (u < 1.0) ? u : (u - 1.0);
}

// y = A x mod m for a 3x3 matrix of residues (row-major).
CASIM_HD void matvec(const uint32_t A[9], const uint32_t x[3], uint32_t y[3], uint32_t c, uint32_t m) {
    uint32_t r[3];
    for (int i = 0; i < 3; ++i) {
        uint64_t acc = (uint64_t)mulmod(A[3 * i], x[0], c, m) + mulmod(A[3 * i + 1], x[1], c, m) +
                       mulmod(A[3 * i + 2], x[2], c, m);
        r[i] = reduce(acc, c, m);
    }
    y[0] = r[0];
    y[1] = r[1];
    y[2] = r[2];
}

CASIM_HD void matmul(const uint32_t A[9], const uint32_t B[9], uint32_t C[9], uint32_t c, uint32_t m) {
    uint32_t R[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            uint64_t acc = (uint64_t)mulmod(A[3 * i], B[j], c, m) + mulmod(A[3 * i + 1], B[3 + j], c, m) +
                           mulmod(A[3 * i + 2], B[6 + j], c, m);
            R[3 * i + j] = reduce(acc, c, m);
        }
    for (int i = 0; i < 9; ++i) C[i] = R[i];
}

// One-step transition matrices with the negative coefficients taken mod m.
inline void base_matrices(uint32_t A1[9], uint32_t A2[9]) {
    const uint32_t a1[9] = {0, 1, 0, 0, 0, 1, M1 - A13N, A12, 0};
    const uint32_t a2[9] = {0, 1, 0, 0, 0, 1, M2 - A23N, 0, A21};
    for (int i = 0; i < 9; ++i) {
        A1[i] = a1[i];
        A2[i] = a2[i];
    }
}

inline void identity(uint32_t I[9]) {
    for (int i = 0; i < 9; ++i) I[i] = (i % 4 == 0) ? 1u : 0u;
}

// B = A^(2^e) mod m.
inline void two_pow(const uint32_t A[9], int e, uint32_t B[9], uint32_t c, uint32_t m) {
    for (int i = 0; i < 9; ++i) B[i] = A[i];
    for (int i = 0; i < e; ++i) matmul(B, B, B, c, m);
}

// B = A^n mod m, n >= 0.
inline void power(const uint32_t A[9], uint64_t n, uint32_t B[9], uint32_t c, uint32_t m) {
    uint32_t W[9];
    for (int i = 0; i < 9; ++i) W[i] = A[i];
    identity(B);
    while (n) {
        if (n & 1) matmul(W, B, B, c, m);
        matmul(W, W, W, c, m);
        n >>= 1;
    }
}

}  // namespace mrg
}  // namespace casim
