// xss_keys.hpp -- key types and the order-preserving key transform.
//
// Every key type of the x86simdsort:: API (u16, i16, _Float16, u32, i32,
// float, u64, i64, double) is handled as its unsigned bit pattern U.  The
// kernels order keys by ord(x), an unsigned integer whose natural order is
// the reference's comparison order (utils/custom-compare.h:11-34,
// src/xss-common-comparators.hpp:37-125):
//
//   unsigned   ord = x
//   signed     ord = x ^ signbit                        (two's complement bias)
//   floating   ord = x ^ (sign ? all-ones : signbit)    (-NaN < -inf < .. < -0 < +0 < .. < +inf < +NaN)
//   hasnan     every NaN (any sign, any payload) -> all-ones, above +inf,
//              so NaNs form one run at the end (NaN is the largest value of
//              the reference's order, utils/custom-compare.h:19-26)
//   descending ord ^= all-ones  (NaN run first, as in xss-common-qsort.h:113-121)
//
// The kernels never store ord: they carry the original bits and recompute
// ord where they need a digit, so kv/select/arg keep NaN payload bits
// (src/xss-common-keyvaluesort.hpp:18-69) and only qsort rewrites NaNs to
// the canonical pattern (src/xss-custom-float.h:34-38,81-85).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace xss {

enum class Kind : int { Unsigned = 0, Signed = 1, Float = 2 };

template <int BITS>
struct UintOf;
template <>
struct UintOf<16> { using type = uint16_t; };
template <>
struct UintOf<32> { using type = uint32_t; };
template <>
struct UintOf<64> { using type = uint64_t; };

// KeyDesc<BITS, KIND>: bit width and ordering kind of a key type.
template <int BITS_, Kind KIND_>
struct KeyDesc {
    static constexpr int BITS = BITS_;
    static constexpr Kind KIND = KIND_;
    using U = typename UintOf<BITS_>::type;
    static constexpr U SIGN = U(U(1) << (BITS_ - 1));
    static constexpr U ONES = U(~U(0));
    // exponent-all-ones pattern (infinity) of the IEEE type of this width
    static constexpr U INF = BITS_ == 16 ? U(0x7c00u)
                            : BITS_ == 32 ? U(0x7f800000u)
                                          : U(0x7ff0000000000000ull);
    // canonical NaN written by qsort(hasnan): quiet_NaN<float|double>(),
    // and 0x7c01 for _Float16 (src/xss-custom-float.h:81-85)
    static constexpr U QNAN = BITS_ == 16 ? U(0x7c01u)
                             : BITS_ == 32 ? U(0x7fc00000u)
                                           : U(0x7ff8000000000000ull);
    static constexpr int PASSES = BITS_ / 8;

    __device__ __forceinline__ static bool is_nan(U x)
    {
        if constexpr (KIND_ == Kind::Float) return U(x & U(~SIGN)) > INF;
        else return false;
    }
    // flip: all-ones for descending, zero for ascending; nan_last: hasnan
    __device__ __forceinline__ static U ord(U x, U flip, bool nan_last)
    {
        U k;
        if constexpr (KIND_ == Kind::Unsigned) {
            k = x;
        } else if constexpr (KIND_ == Kind::Signed) {
            k = U(x ^ SIGN);
        } else {
            using S = typename std::make_signed<U>::type;
            U m = U(U(S(x) >> (BITS_ - 1)) | SIGN);
            k = U(x ^ m);
            if (nan_last && is_nan(x)) k = ONES;
        }
        return U(k ^ flip);
    }
    __device__ __forceinline__ static uint32_t digit(U x, U flip, bool nan_last, int shift)
    {
        return uint32_t(ord(x, flip, nan_last) >> shift) & 0xffu;
    }
};

}  // namespace xss
