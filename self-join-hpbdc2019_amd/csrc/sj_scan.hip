// Exclusive scans with the total appended (out[n] = sum of in[0..n)), as a
// single-pass decoupled look-back scan from rocPRIM: one launch over the
// data instead of a reduce / block-sums / apply sequence.  Used for the
// stream-ordered scans of grid construction and joins; the many-join graph
// keeps its caller-buffered three-phase scan (sj_prims.hip).
#include <rocprim/device/device_scan.hpp>
#include <rocprim/iterator/counting_iterator.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

#include "sj_common.hpp"

namespace sj {

namespace {

// element i of the scan input: in[i] below n, 0 at n (so out[n] is the total)
template <class Tacc>
struct PaddedInput {
    const uint32_t *in;
    uint64_t n;
    __host__ __device__ Tacc operator()(size_t i) const { return i < n ? (Tacc)in[i] : (Tacc)0; }
};

template <class Tacc>
sj_status scan_lookback(Pool *pool, hipStream_t s, const uint32_t *in, uint64_t n, Tacc *out) {
    auto input = rocprim::make_transform_iterator(rocprim::counting_iterator<size_t>(0), PaddedInput<Tacc>{in, n});
    size_t bytes = 0;
    SJ_TRY_HIP(rocprim::exclusive_scan(nullptr, bytes, input, out, (Tacc)0, (size_t)(n + 1), rocprim::plus<Tacc>(), s));
    DBuf<unsigned char> tmp;
    SJ_TRY(tmp.alloc(pool, bytes ? bytes : 1));
    SJ_TRY_HIP(rocprim::exclusive_scan(tmp.p, bytes, input, out, (Tacc)0, (size_t)(n + 1), rocprim::plus<Tacc>(), s));
    return SJ_OK;
}

// element i: a[i] + b[i] below n, 0 at n
struct PaddedSum {
    const uint32_t *a, *b;
    uint64_t n;
    __host__ __device__ uint64_t operator()(size_t i) const { return i < n ? (uint64_t)a[i] + b[i] : 0ull; }
};

}  // namespace

sj_status scan_exclusive_sum2_u32_to_u64(Pool *pool, hipStream_t s, const uint32_t *a, const uint32_t *b,
                                         uint64_t n, uint64_t *out) {
    auto input = rocprim::make_transform_iterator(rocprim::counting_iterator<size_t>(0), PaddedSum{a, b, n});
    size_t bytes = 0;
    SJ_TRY_HIP(rocprim::exclusive_scan(nullptr, bytes, input, out, (uint64_t)0, (size_t)(n + 1),
                                       rocprim::plus<uint64_t>(), s));
    DBuf<unsigned char> tmp;
    SJ_TRY(tmp.alloc(pool, bytes ? bytes : 1));
    SJ_TRY_HIP(rocprim::exclusive_scan(tmp.p, bytes, input, out, (uint64_t)0, (size_t)(n + 1),
                                       rocprim::plus<uint64_t>(), s));
    return SJ_OK;
}

sj_status scan_exclusive_u32_to_u64(Pool *pool, hipStream_t s, const uint32_t *in, uint64_t n, uint64_t *out) {
    return scan_lookback<uint64_t>(pool, s, in, n, out);
}

sj_status scan_exclusive_u32_to_u32(Pool *pool, hipStream_t s, const uint32_t *in, uint64_t n, uint32_t *out) {
    return scan_lookback<uint32_t>(pool, s, in, n, out);
}

}  // namespace sj
