// exp and log in binary32 arithmetic with float-float internals, for the
// kernels whose results follow the reference's FP32 build (glibc expf/logf,
// which round within about 0.5 ulp): the results here are the correctly
// rounded values but for arguments whose exact value lies within ~2^-40
// relative of a rounding boundary, and no binary64 arithmetic is used.
//
// Method (Tang, "Table-driven implementation of the exponential function in
// IEEE floating-point arithmetic", ACM TOMS 15(2), 1989, and its logarithm
// companion, ACM TOMS 16(4), 1990), carried out in float-float with FMA
// for the exact products (Dekker/Knuth two-sum and two-product):
//   exp: n = rint(x 32/ln2), x = (32k + j) ln2/32 + r, |r| <= ln2/64, r in
//        float-float from a three-part ln2/32 (the first part's 9 bits make
//        n*C1 exact), exp(x) = 2^k * 2^(j/32) * (1 + r + r^2/2 + ... + r^5/120);
//   log: x = 2^e m, m in [sqrt(1/2), sqrt(2)), m = c_i (1 + f) with c_i the
//        centre of m's 1/256 (1/128 above 1) interval, c = 1 at the two
//        intervals beside 1; log x = e ln2 - log(1/c_i) + log1p(f), log1p
//        to f^6, f = m/c_i - 1 in float-float from an exact product.
// Tables: 2^(j/32) and -log(1/c_i) as hi + lo, 1/c_i rounded to binary32,
// computed at 200 bits (mpmath) and rounded.
#pragma once
#include <hip/hip_runtime.h>

namespace cress {
namespace xm {

__host__ __device__ __forceinline__ void two_sum(float a, float b, float &s, float &e) {
  s = a + b;
  const float bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

__host__ __device__ __forceinline__ float ldexp2(float y, int k) {
  // y * 2^k in two exact power-of-two factors (one rounding at most)
  const int k1 = k >> 1, k2 = k - k1;
  return (y * __builtin_bit_cast(float, (k1 + 127) << 23)) * __builtin_bit_cast(float, (k2 + 127) << 23);
}

struct FF { float hi, lo; };
struct LogRow { float inv, lh, ll; };

__host__ __device__ __forceinline__ const FF *exp_tab() {
  static constexpr FF t[32] = {
    {0x1.0000000000000p+0f, 0x0.0p+0f},
    {0x1.059b0e0000000p+0f, -0x1.9d4f520000000p-25f},
    {0x1.0b55860000000p+0f, 0x1.9f31220000000p-25f},
    {0x1.11301e0000000p+0f, -0x1.fdb4960000000p-25f},
    {0x1.172b840000000p+0f, -0x1.c157420000000p-27f},
    {0x1.1d48740000000p+0f, -0x1.d2e8ca0000000p-25f},
    {0x1.2387a60000000p+0f, 0x1.ceac480000000p-25f},
    {0x1.29e9e00000000p+0f, -0x1.5c04240000000p-25f},
    {0x1.306fe00000000p+0f, 0x1.4636e20000000p-25f},
    {0x1.371a740000000p+0f, -0x1.18aac60000000p-25f},
    {0x1.3dea640000000p+0f, 0x1.8246840000000p-25f},
    {0x1.44e0860000000p+0f, 0x1.8624b40000000p-30f},
    {0x1.4bfdae0000000p+0f, -0x1.593abc0000000p-25f},
    {0x1.5342b60000000p+0f, -0x1.2c56100000000p-25f},
    {0x1.5ab07e0000000p+0f, -0x1.5bd5ec0000000p-27f},
    {0x1.6247ec0000000p+0f, -0x1.f8b5500000000p-25f},
    {0x1.6a09e60000000p+0f, 0x1.9fcef40000000p-26f},
    {0x1.71f75e0000000p+0f, 0x1.1d8bee0000000p-25f},
    {0x1.7a11480000000p+0f, -0x1.829fd00000000p-25f},
    {0x1.82589a0000000p+0f, -0x1.accc7c0000000p-26f},
    {0x1.8ace540000000p+0f, 0x1.15506e0000000p-27f},
    {0x1.93737c0000000p+0f, -0x1.e647440000000p-25f},
    {0x1.9c49180000000p+0f, 0x1.51f8480000000p-27f},
    {0x1.a5503c0000000p+0f, -0x1.b83b540000000p-25f},
    {0x1.ae89fa0000000p+0f, -0x1.a94b140000000p-26f},
    {0x1.b7f7700000000p+0f, -0x1.a094380000000p-25f},
    {0x1.c199be0000000p+0f, -0x1.3d56b20000000p-27f},
    {0x1.cb720e0000000p+0f, -0x1.8837cc0000000p-27f},
    {0x1.d5818e0000000p+0f, -0x1.822dbc0000000p-27f},
    {0x1.dfc9740000000p+0f, -0x1.908c940000000p-25f},
    {0x1.ea4afa0000000p+0f, 0x1.52486c0000000p-27f},
    {0x1.f507660000000p+0f, -0x1.246eb00000000p-26f},
  };
  return t;
}

__host__ __device__ __forceinline__ const LogRow *log_tab() {
  static constexpr LogRow t[131] = {
    {0x1.6b0f960000000p+0f, -0x1.65c74e0000000p-2f, 0x1.bec4f00000000p-29f},
    {0x1.6a0eda0000000p+0f, -0x1.62f2320000000p-2f, 0x1.b9a7320000000p-27f},
    {0x1.690f880000000p+0f, -0x1.601f140000000p-2f, -0x1.1ac6f40000000p-28f},
    {0x1.6719f40000000p+0f, -0x1.5a8cb00000000p-2f, 0x1.f0481a0000000p-28f},
    {0x1.6524f80000000p+0f, -0x1.54f4300000000p-2f, -0x1.8f7c340000000p-27f},
    {0x1.63356c0000000p+0f, -0x1.4f63800000000p-2f, -0x1.3a980c0000000p-30f},
    {0x1.614b360000000p+0f, -0x1.49da7e0000000p-2f, 0x1.00cefa0000000p-28f},
    {0x1.5f66440000000p+0f, -0x1.4459200000000p-2f, -0x1.69cfa00000000p-29f},
    {0x1.5d867c0000000p+0f, -0x1.3edf460000000p-2f, 0x1.efa5f00000000p-28f},
    {0x1.5babcc0000000p+0f, -0x1.396ce20000000p-2f, -0x1.8ddfa80000000p-29f},
    {0x1.59d6200000000p+0f, -0x1.3401e40000000p-2f, 0x1.51346e0000000p-30f},
    {0x1.5805600000000p+0f, -0x1.2e9e2c0000000p-2f, 0x1.c7b75e0000000p-28f},
    {0x1.56397c0000000p+0f, -0x1.2941b00000000p-2f, -0x1.730d700000000p-27f},
    {0x1.54725e0000000p+0f, -0x1.23ec580000000p-2f, -0x1.37ae920000000p-28f},
    {0x1.52aff60000000p+0f, -0x1.1e9e180000000p-2f, -0x1.e44cf80000000p-29f},
    {0x1.50f22e0000000p+0f, -0x1.1956d40000000p-2f, 0x1.e90f420000000p-28f},
    {0x1.4f38f60000000p+0f, -0x1.14167e0000000p-2f, -0x1.9d9de00000000p-28f},
    {0x1.4d843c0000000p+0f, -0x1.0edd060000000p-2f, -0x1.0de0200000000p-28f},
    {0x1.4bd3ee0000000p+0f, -0x1.09aa580000000p-2f, 0x1.764e4c0000000p-28f},
    {0x1.4a27fa0000000p+0f, -0x1.047e5e0000000p-2f, -0x1.8f41d60000000p-29f},
    {0x1.4880520000000p+0f, -0x1.feb2220000000p-3f, -0x1.da81f40000000p-29f},
    {0x1.46dce40000000p+0f, -0x1.f474b60000000p-3f, 0x1.d906f60000000p-30f},
    {0x1.453d9e0000000p+0f, -0x1.ea44480000000p-3f, -0x1.b0955e0000000p-28f},
    {0x1.43a2740000000p+0f, -0x1.e020d20000000p-3f, -0x1.c8d6a40000000p-29f},
    {0x1.420b520000000p+0f, -0x1.d60a160000000p-3f, 0x1.1df95e0000000p-28f},
    {0x1.40782e0000000p+0f, -0x1.cc00120000000p-3f, -0x1.2b67860000000p-28f},
    {0x1.3ee8f40000000p+0f, -0x1.c2028a0000000p-3f, 0x1.7a01940000000p-29f},
    {0x1.3d5d9a0000000p+0f, -0x1.b811780000000p-3f, -0x1.a704760000000p-28f},
    {0x1.3bd60e0000000p+0f, -0x1.ae2caa0000000p-3f, 0x1.06350c0000000p-29f},
    {0x1.3a52440000000p+0f, -0x1.a4540c0000000p-3f, 0x1.832aa00000000p-28f},
    {0x1.38d22e0000000p+0f, -0x1.9a877e0000000p-3f, -0x1.aea8780000000p-33f},
    {0x1.3755be0000000p+0f, -0x1.90c6e20000000p-3f, 0x1.10686a0000000p-28f},
    {0x1.35dce60000000p+0f, -0x1.8712140000000p-3f, 0x1.8bc59a0000000p-29f},
    {0x1.34679a0000000p+0f, -0x1.7d68fe0000000p-3f, -0x1.cbd6ac0000000p-29f},
    {0x1.32f5ce0000000p+0f, -0x1.73cb8a0000000p-3f, -0x1.b9fa260000000p-28f},
    {0x1.3187760000000p+0f, -0x1.6a39a00000000p-3f, -0x1.477a700000000p-28f},
    {0x1.301c820000000p+0f, -0x1.60b30c0000000p-3f, 0x1.f3dae80000000p-29f},
    {0x1.2eb4ea0000000p+0f, -0x1.5737cc0000000p-3f, 0x1.1f9cc80000000p-29f},
    {0x1.2d50a00000000p+0f, -0x1.4dc7b80000000p-3f, -0x1.7bc1c80000000p-31f},
    {0x1.2bef980000000p+0f, -0x1.4462b40000000p-3f, 0x1.0d93120000000p-29f},
    {0x1.2a91ca0000000p+0f, -0x1.3b08bc0000000p-3f, -0x1.afe5140000000p-32f},
    {0x1.2937260000000p+0f, -0x1.31b9960000000p-3f, -0x1.5749f00000000p-28f},
    {0x1.27dfa40000000p+0f, -0x1.28753e0000000p-3f, -0x1.e235740000000p-28f},
    {0x1.268b380000000p+0f, -0x1.1f3b940000000p-3f, 0x1.0368b00000000p-29f},
    {0x1.2539d80000000p+0f, -0x1.160c800000000p-3f, -0x1.8964f60000000p-28f},
    {0x1.23eb7a0000000p+0f, -0x1.0ce7f00000000p-3f, -0x1.8998500000000p-28f},
    {0x1.22a0120000000p+0f, -0x1.03cdc00000000p-3f, 0x1.05c27e0000000p-28f},
    {0x1.2157980000000p+0f, -0x1.f57bc80000000p-4f, 0x1.9bfe8a0000000p-30f},
    {0x1.2012020000000p+0f, -0x1.e3708c0000000p-4f, 0x1.59f6fa0000000p-29f},
    {0x1.1ecf440000000p+0f, -0x1.d1797c0000000p-4f, 0x1.779b280000000p-30f},
    {0x1.1d8f560000000p+0f, -0x1.bf96800000000p-4f, -0x1.f3f9400000000p-29f},
    {0x1.1c52300000000p+0f, -0x1.adc7820000000p-4f, -0x1.96baa20000000p-30f},
    {0x1.1b17c60000000p+0f, -0x1.9c0c2c0000000p-4f, 0x1.6cb6b40000000p-30f},
    {0x1.19e0120000000p+0f, -0x1.8a647e0000000p-4f, 0x1.8dc47c0000000p-29f},
    {0x1.18ab080000000p+0f, -0x1.78d0200000000p-4f, 0x1.b84fa60000000p-29f},
    {0x1.1778a20000000p+0f, -0x1.674f0e0000000p-4f, -0x1.c6cb4c0000000p-29f},
    {0x1.1648d60000000p+0f, -0x1.55e10e0000000p-4f, -0x1.0701920000000p-31f},
    {0x1.151b9a0000000p+0f, -0x1.4485dc0000000p-4f, -0x1.1b7bf40000000p-29f},
    {0x1.13f0e80000000p+0f, -0x1.333d740000000p-4f, 0x1.7cf8200000000p-29f},
    {0x1.12c8b80000000p+0f, -0x1.2207ac0000000p-4f, -0x1.0f0a8e0000000p-29f},
    {0x1.11a3020000000p+0f, -0x1.10e4620000000p-4f, 0x1.a6a2fc0000000p-29f},
    {0x1.107fbc0000000p+0f, -0x1.ffa6940000000p-5f, -0x1.b572600000000p-30f},
    {0x1.0f5ee00000000p+0f, -0x1.dda8b80000000p-5f, 0x1.cc08e60000000p-32f},
    {0x1.0e40660000000p+0f, -0x1.bbced40000000p-5f, 0x1.65c3100000000p-31f},
    {0x1.0d24460000000p+0f, -0x1.9a188e0000000p-5f, 0x1.1843b60000000p-34f},
    {0x1.0c0a780000000p+0f, -0x1.78858a0000000p-5f, 0x1.b9510e0000000p-30f},
    {0x1.0af2f80000000p+0f, -0x1.5715e00000000p-5f, 0x1.7f86380000000p-30f},
    {0x1.09ddba0000000p+0f, -0x1.35c8b20000000p-5f, -0x1.9426080000000p-30f},
    {0x1.08cabc0000000p+0f, -0x1.149e560000000p-5f, -0x1.000b400000000p-30f},
    {0x1.07b9f20000000p+0f, -0x1.e72bcc0000000p-6f, -0x1.8279b40000000p-31f},
    {0x1.06ab5a0000000p+0f, -0x1.a55f620000000p-6f, -0x1.31710a0000000p-32f},
    {0x1.059eea0000000p+0f, -0x1.63d6160000000p-6f, 0x1.cb7a160000000p-33f},
    {0x1.04949c0000000p+0f, -0x1.228f820000000p-6f, -0x1.fa8b440000000p-32f},
    {0x1.038c6c0000000p+0f, -0x1.c3177c0000000p-7f, 0x1.6714420000000p-32f},
    {0x1.0286500000000p+0f, -0x1.4192bc0000000p-7f, 0x1.a5f3500000000p-33f},
    {0x1.0182440000000p+0f, -0x1.8121bc0000000p-8f, 0x1.74f2f20000000p-33f},
    {0x1.0000000000000p+0f, 0x0.0p+0f, 0x0.0p+0f},
    {0x1.0000000000000p+0f, 0x0.0p+0f, 0x0.0p+0f},
    {0x1.fa11ca0000000p-1f, 0x1.7dc49e0000000p-7f, 0x1.e042b80000000p-33f},
    {0x1.f6310a0000000p-1f, 0x1.3cea5e0000000p-6f, -0x1.72b4700000000p-35f},
    {0x1.f25f640000000p-1f, 0x1.b9fc0a0000000p-6f, 0x1.f5f2340000000p-31f},
    {0x1.ee9c800000000p-1f, 0x1.1b0d900000000p-5f, 0x1.247b320000000p-30f},
    {0x1.eae8080000000p-1f, 0x1.58a5b60000000p-5f, -0x1.06e3640000000p-30f},
    {0x1.e741aa0000000p-1f, 0x1.95c8360000000p-5f, 0x1.991c7e0000000p-30f},
    {0x1.e3a9180000000p-1f, 0x1.d276b20000000p-5f, 0x1.6d85ae0000000p-32f},
    {0x1.e01e020000000p-1f, 0x1.0759820000000p-4f, 0x1.26391c0000000p-30f},
    {0x1.dca01e0000000p-1f, 0x1.253f620000000p-4f, -0x1.bebd7c0000000p-29f},
    {0x1.d92f220000000p-1f, 0x1.42edce0000000p-4f, -0x1.966e440000000p-30f},
    {0x1.d5cac80000000p-1f, 0x1.60658a0000000p-4f, 0x1.a6ea180000000p-29f},
    {0x1.d272ca0000000p-1f, 0x1.7da76a0000000p-4f, -0x1.f09da60000000p-29f},
    {0x1.cf26e60000000p-1f, 0x1.9ab4220000000p-4f, 0x1.480cec0000000p-30f},
    {0x1.cbe6da0000000p-1f, 0x1.b78c7e0000000p-4f, -0x1.a9e24a0000000p-29f},
    {0x1.c8b2660000000p-1f, 0x1.d4313a0000000p-4f, 0x1.2d966c0000000p-29f},
    {0x1.c5894e0000000p-1f, 0x1.f0a3040000000p-4f, -0x1.bba74e0000000p-30f},
    {0x1.c26b540000000p-1f, 0x1.0671500000000p-3f, -0x1.86b4d20000000p-28f},
    {0x1.bf583e0000000p-1f, 0x1.14785c0000000p-3f, 0x1.b9d0b00000000p-29f},
    {0x1.bc4fd60000000p-1f, 0x1.2266f40000000p-3f, -0x1.aeb4a60000000p-28f},
    {0x1.b951e20000000p-1f, 0x1.303d740000000p-3f, 0x1.8c8ffc0000000p-28f},
    {0x1.b65e2e0000000p-1f, 0x1.3dfc2c0000000p-3f, 0x1.3663160000000p-30f},
    {0x1.b374840000000p-1f, 0x1.4ba3720000000p-3f, 0x1.a6957c0000000p-29f},
    {0x1.b094b40000000p-1f, 0x1.59338a0000000p-3f, -0x1.2cfbee0000000p-28f},
    {0x1.adbe880000000p-1f, 0x1.66acd40000000p-3f, 0x1.cab5440000000p-33f},
    {0x1.aaf1d20000000p-1f, 0x1.740f940000000p-3f, -0x1.fe422c0000000p-34f},
    {0x1.a82e660000000p-1f, 0x1.815c060000000p-3f, -0x1.8f2a000000000p-29f},
    {0x1.a574100000000p-1f, 0x1.8e92900000000p-3f, 0x1.4436a20000000p-30f},
    {0x1.a2c2a80000000p-1f, 0x1.9bb3660000000p-3f, -0x1.70408e0000000p-28f},
    {0x1.a01a020000000p-1f, 0x1.a8bece0000000p-3f, -0x1.efa1c80000000p-32f},
    {0x1.9d79f20000000p-1f, 0x1.b5b5180000000p-3f, -0x1.7e094a0000000p-28f},
    {0x1.9ae24e0000000p-1f, 0x1.c296880000000p-3f, 0x1.21831a0000000p-28f},
    {0x1.9852f00000000p-1f, 0x1.cf635a0000000p-3f, -0x1.bec7420000000p-28f},
    {0x1.95cbb00000000p-1f, 0x1.dc1bce0000000p-3f, -0x1.aa09ba0000000p-30f},
    {0x1.934c680000000p-1f, 0x1.e8c0260000000p-3f, -0x1.eab4b40000000p-28f},
    {0x1.90d4f20000000p-1f, 0x1.f550a00000000p-3f, 0x1.d96f6a0000000p-28f},
    {0x1.8e65280000000p-1f, 0x1.00e6c40000000p-2f, -0x1.d4abf80000000p-28f},
    {0x1.8bfce80000000p-1f, 0x1.071b860000000p-2f, 0x1.9ab21a0000000p-31f},
    {0x1.899c100000000p-1f, 0x1.0d46b40000000p-2f, -0x1.32a4580000000p-29f},
    {0x1.87427c0000000p-1f, 0x1.1368700000000p-2f, -0x1.7b15d40000000p-28f},
    {0x1.84f00c0000000p-1f, 0x1.1980d40000000p-2f, -0x1.757b920000000p-27f},
    {0x1.82a4a00000000p-1f, 0x1.1f8ffa0000000p-2f, 0x1.24517a0000000p-29f},
    {0x1.8060180000000p-1f, 0x1.2596020000000p-2f, -0x1.c411380000000p-27f},
    {0x1.7e22560000000p-1f, 0x1.2b93020000000p-2f, -0x1.90ec5a0000000p-27f},
    {0x1.7beb3a0000000p-1f, 0x1.31871a0000000p-2f, 0x1.0510640000000p-28f},
    {0x1.79baa60000000p-1f, 0x1.3772680000000p-2f, 0x1.3fec320000000p-29f},
    {0x1.7790820000000p-1f, 0x1.3d54f80000000p-2f, -0x1.7e08e40000000p-30f},
    {0x1.756cac0000000p-1f, 0x1.432ef20000000p-2f, 0x1.f09d020000000p-27f},
    {0x1.734f0c0000000p-1f, 0x1.4900680000000p-2f, 0x1.d8013a0000000p-27f},
    {0x1.7137860000000p-1f, 0x1.4ec9760000000p-2f, -0x1.f7ff620000000p-28f},
    {0x1.6f26020000000p-1f, 0x1.548a2a0000000p-2f, 0x1.4dba4c0000000p-27f},
    {0x1.6d1a620000000p-1f, 0x1.5a42ac0000000p-2f, 0x1.9a67f20000000p-29f},
    {0x1.6b14900000000p-1f, 0x1.5ff3080000000p-2f, 0x1.d4f27c0000000p-27f},
    {0x1.6a0eda0000000p-1f, 0x1.62d62e0000000p-2f, 0x1.7835ae0000000p-27f},
  };
  return t;
}

}  // namespace xm

__host__ __device__ __forceinline__ float xexpf(float x) {
  using namespace xm;
  if (x != x) return x + x;
  if (x > 88.8f) return __builtin_huge_valf();
  if (x < -104.0f) return 0.0f;
  const float n = __builtin_rintf(x * 0x1.715476p+5f);            // 32/ln2
  const int ni = (int)n;
  const int k = ni >> 5, j = ni & 31;
  // r = x - n ln2/32, ln2/32 = C1 + C2 + C3
  const float C1 = 0x1.63p-6f, C2 = -0x1.bd0106p-18f, C3 = 0x1.cf79acp-45f;
  const float rh0 = __builtin_fmaf(-n, C1, x);                     // exact
  const float p = n * C2, pe = __builtin_fmaf(n, C2, -p);
  float s, es;
  two_sum(rh0, -p, s, es);
  const float rl0 = (es - pe) - n * C3;
  const float rh = s + rl0, rl = rl0 - (rh - s);
  // exp(r) - 1 = rh + q + tail, q = rh^2/2 from an exact square
  const float s2 = rh * rh, s2e = __builtin_fmaf(rh, rh, -s2);
  const float q = 0.5f * s2;
  const float tail = (rl + rh * rl) + (0.5f * s2e + rh * s2 * (0x1.555556p-3f + rh * (0x1.555556p-5f + rh * 0x1.111112p-7f)));
  const FF T = exp_tab()[j];
  const float a = T.hi * rh, ae = __builtin_fmaf(T.hi, rh, -a);
  const float b = T.hi * q, be = __builtin_fmaf(T.hi, q, -b);
  const float small = ((ae + be) + T.hi * tail) + (T.lo + T.lo * (rh + q));
  float h1, e1, h2, e2;
  two_sum(T.hi, a, h1, e1);
  two_sum(h1, b, h2, e2);
  const float y = h2 + ((e1 + e2) + small);
  return ldexp2(y, k);
}

__host__ __device__ __forceinline__ float xlogf(float x) {
  using namespace xm;
  if (!(x > 0.0f)) return x == 0.0f ? -__builtin_huge_valf() : __builtin_nanf("");
  if (x == __builtin_huge_valf()) return x;
  unsigned u = __builtin_bit_cast(unsigned, x);
  int e = 0;
  if (u < 0x00800000u) {                  // subnormal: scale into the normal range
    u = __builtin_bit_cast(unsigned, x * 0x1p23f);
    e = -23;
  }
  // m in [sqrt(1/2), sqrt(2)): m = x 2^-e
  const unsigned t = u - 0x3f3504f3u;
  const int eb = (int)t >> 23;
  e += eb;
  const unsigned mb = u - ((unsigned)eb << 23);
  const float m = __builtin_bit_cast(float, mb);
  const LogRow R = log_tab()[(mb - 0x3f330000u) >> 16];
  // f = m/c - 1 = fh + fl exactly: m*inv as an exact product, ph - 1 exact
  const float ph = m * R.inv, pl = __builtin_fmaf(m, R.inv, -ph);
  const float fh = ph - 1.0f, fl = pl;
  const float f2 = fh * fh, f2e = __builtin_fmaf(fh, fh, -f2);
  const float poly = fh * f2 * (0x1.555556p-2f - fh * (0.25f - fh * (0x1.99999ap-3f - fh * 0x1.555556p-3f)));
  const float ef = (float)e;
  const float LH = 0x1.62e4p-1f, LM = 0x1.7f7d1cp-20f, LL = 0x1.ef357ap-45f;
  const float pm = ef * LM, pme = __builtin_fmaf(ef, LM, -pm);
  const float E = (((pme + ef * LL) + R.ll) + (poly - 0.5f * f2e)) + (fl - fh * fl);
  float s1, t1, s2, t2, s3, t3, s4, t4;
  two_sum(ef * LH, R.lh, s1, t1);
  two_sum(s1, pm, s2, t2);
  two_sum(s2, fh, s3, t3);
  two_sum(s3, -0.5f * f2, s4, t4);
  return s4 + (((t1 + t2) + (t3 + t4)) + E);
}

}  // namespace cress
