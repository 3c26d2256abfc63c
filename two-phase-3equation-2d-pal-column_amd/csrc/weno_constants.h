// The reference expressions of include/Utility.hpp:12-39 and 64-90 evaluated in IEEE
// double, written as hexfloats so that the device compiler needs no library call.
// tests/test_weno_constants.py evaluates those expressions again and compares bit for bit,
// and checks the tables against polynomial exactness at the Gauss-Lobatto nodes.
#pragma once
namespace wfv {
// linear weights WENO_D[gp][k]
constexpr double WENO_D[4][3] = {
    {0x1.3333333333333p-2, 0x1.3333333333333p-1, 0x1.999999999999ap-4},
    {0x1.49e489091c28dp-3, 0x1.2c37dac37dac3p-1, 0x1.029e05f476933p-2},
    {0x1.029e05f476933p-2, 0x1.2c37dac37dac3p-1, 0x1.49e489091c28dp-3},
    {0x1.999999999999ap-4, 0x1.3333333333333p-1, 0x1.3333333333333p-2},
};
// candidate coefficients WENO_H[gp][k][m]: h_k = (c0*s[k] + c1*s[k+1]) + c2*s[k+2]
constexpr double WENO_H[4][3][3] = {
    {{-0x1.5555555555555p-3, 0x1.aaaaaaaaaaaabp-1, 0x1.5555555555555p-2}, {0x1.5555555555555p-2, 0x1.aaaaaaaaaaaabp-1, -0x1.5555555555555p-3}, {0x1.d555555555555p+0, -0x1.2aaaaaaaaaaabp+0, 0x1.5555555555555p-2}},
    {{-0x1.071b50502190fp-3, 0x1.ec147e7e20ffcp-2, 0x1.4bbc94d4f7e46p-1}, {0x1.85ae1817ba996p-4, 0x1.0888888888889p+0, -0x1.071b50502190fp-3}, {0x1.51992d0cfb854p+0, -0x1.a7d03a39dcbb8p-2, 0x1.85ae1817ba996p-4}},
    {{0x1.85ae1817ba996p-4, -0x1.a7d03a39dcbb8p-2, 0x1.51992d0cfb854p+0}, {-0x1.071b50502190fp-3, 0x1.0888888888889p+0, 0x1.85ae1817ba996p-4}, {0x1.4bbc94d4f7e46p-1, 0x1.ec147e7e20ffcp-2, -0x1.071b50502190fp-3}},
    {{0x1.5555555555555p-2, -0x1.2aaaaaaaaaaabp+0, 0x1.d555555555555p+0}, {-0x1.5555555555555p-3, 0x1.aaaaaaaaaaaabp-1, 0x1.5555555555555p-2}, {0x1.5555555555555p-2, 0x1.aaaaaaaaaaaabp-1, -0x1.5555555555555p-3}},
};
}  // namespace wfv
