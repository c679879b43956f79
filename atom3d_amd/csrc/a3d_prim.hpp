// Device-wide sorts, scans and selections come from rocPRIM, called
// directly: rocprim::radix_sort_pairs / radix_sort_keys / select /
// exclusive_scan.  This header adds the one shorthand the library uses
// everywhere, the exclusive prefix sum.
#pragma once

#include <cstring>  // rocPRIM's texture iterator calls memset on the host

#include <rocprim/rocprim.hpp>

namespace a3d {

// out[i] = in[0] + ... + in[i - 1] (out[0] = 0), in the output's type; the
// usual two calls: temp == nullptr sets bytes, then the scan.
template <typename In, typename Out>
inline hipError_t exclusive_sum(void* temp, size_t& bytes, In in, Out* out, size_t n, hipStream_t stream = 0) {
  return rocprim::exclusive_scan(temp, bytes, in, out, Out(0), n, rocprim::plus<Out>(), stream);
}

}  // namespace a3d
