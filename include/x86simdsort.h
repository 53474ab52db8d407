/*
 * x86simdsort.h -- drop-in C++ API of the MI355X sort/select library.
 *
 * The declarations are those of the reference's public header
 * (lib/x86simdsort.h:15-70): the same templates, argument order and default
 * arguments, so the exported specialisations carry the same mangled names
 * (e.g. _ZN11x86simdsort5qsortIfEEvPT_mbb) and a program built against the
 * reference links against libxss_amd.so unchanged.  These are host-pointer
 * calls: each one copies the array to the GPU, sorts there and copies the
 * result back before returning (synchronous, like the reference).  A failed
 * GPU call throws std::runtime_error; results are never silently wrong.
 * Device-pointer, stream-ordered forms are the C ABI of xss_gpu.h.
 *
 * Types: qsort/qselect/partial_qsort/argsort/argselect for uint16_t,
 * int16_t, _Float16 (where the compiler has it), float, double, uint32_t,
 * int32_t, uint64_t, int64_t; keyvalue_* for the 36 pairs of
 * {uint32_t, int32_t, float, uint64_t, int64_t, double}, and for uint16_t,
 * int16_t and _Float16 keys with those six value types (an addition: the
 * reference has no 16-bit key-value sort).
 */
#ifndef XSS_AMD_X86SIMDSORT_H
#define XSS_AMD_X86SIMDSORT_H

#include <stdint.h>

#include <cstddef>
#include <functional>
#include <numeric>
#include <stdexcept>
#include <type_traits>
#include <vector>

#include "xss_gpu.h"

#define XSS_EXPORT_SYMBOL __attribute__((visibility("default")))

namespace x86simdsort {

template <typename T>
XSS_EXPORT_SYMBOL void qsort(T *arr, size_t arrsize, bool hasnan = false, bool descending = false);

template <typename T>
XSS_EXPORT_SYMBOL void qselect(T *arr, size_t k, size_t arrsize, bool hasnan = false, bool descending = false);

template <typename T>
XSS_EXPORT_SYMBOL void partial_qsort(T *arr, size_t k, size_t arrsize, bool hasnan = false, bool descending = false);

template <typename T>
XSS_EXPORT_SYMBOL std::vector<size_t> argsort(T *arr, size_t arrsize, bool hasnan = false, bool descending = false);

template <typename T>
XSS_EXPORT_SYMBOL std::vector<size_t> argselect(T *arr, size_t k, size_t arrsize, bool hasnan = false);

template <typename T1, typename T2>
XSS_EXPORT_SYMBOL void keyvalue_qsort(T1 *key, T2 *val, size_t arrsize, bool hasnan = false,
                                      bool descending = false);

template <typename T1, typename T2>
XSS_EXPORT_SYMBOL void keyvalue_select(T1 *key, T2 *val, size_t k, size_t arrsize, bool hasnan = false,
                                       bool descending = false);

template <typename T1, typename T2>
XSS_EXPORT_SYMBOL void keyvalue_partial_sort(T1 *key, T2 *val, size_t k, size_t arrsize, bool hasnan = false,
                                             bool descending = false);

namespace detail {
// Sorts `n` objects of `obj_bytes` each (host memory) by the 32/64-bit keys
// in `keys` (host memory), ascending, hasnan = false, on the GPU.  Defined in
// the library for KeyT in {float, double, uint32_t, int32_t, uint64_t, int64_t}.
template <typename KeyT>
XSS_EXPORT_SYMBOL void object_sort_by_keys(void *objs, size_t obj_bytes, const KeyT *keys, size_t n);
}  // namespace detail

/*
 * object_qsort (lib/x86simdsort.h:72-106): sort an array of objects by a key
 * computed from each object.  As in the reference the keys are computed on
 * the host by key_func (so any callable works and the key is exactly what
 * the caller's compiler computes), U must be a 32- or 64-bit integral type
 * and the key 32 or 64 bits wide.  The (key, index) sort and the permutation
 * of the objects run on the GPU (an out-of-place gather instead of the
 * reference's in-place cycle walk); the reference sorts with hasnan = false,
 * ascending, and so does this.  T must be trivially copyable.
 */
template <typename T, typename U, typename Func>
XSS_EXPORT_SYMBOL void object_qsort(T *arr, U arrsize, Func key_func)
{
    static_assert(std::is_integral<U>::value, "arrsize must be an integral type");
    static_assert(sizeof(U) == sizeof(int32_t) || sizeof(U) == sizeof(int64_t), "arrsize must be 32 or 64 bits");
    using key_t = typename decltype(std::function{key_func})::result_type;
    static_assert(sizeof(key_t) == sizeof(int32_t) || sizeof(key_t) == sizeof(int64_t),
                  "key_func return type must be 32 or 64 bits");
    static_assert(std::is_trivially_copyable<T>::value, "objects are moved as bytes on the GPU");
    // the library's key type of the same width, signedness and kind
    using lib_key_t = typename std::conditional<
            std::is_floating_point<key_t>::value,
            typename std::conditional<sizeof(key_t) == 4, float, double>::type,
            typename std::conditional<
                    std::is_signed<key_t>::value,
                    typename std::conditional<sizeof(key_t) == 4, int32_t, int64_t>::type,
                    typename std::conditional<sizeof(key_t) == 4, uint32_t, uint64_t>::type>::type>::type;
    if (arrsize <= 1) return;
    std::vector<lib_key_t> keys(static_cast<size_t>(arrsize));
    for (U ii = 0; ii < arrsize; ++ii) keys[ii] = static_cast<lib_key_t>(key_func(arr[ii]));
    detail::object_sort_by_keys<lib_key_t>(static_cast<void *>(arr), sizeof(T), keys.data(),
                                           static_cast<size_t>(arrsize));
}

}  // namespace x86simdsort

#endif /* XSS_AMD_X86SIMDSORT_H */
