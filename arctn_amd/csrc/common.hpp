// Shared definitions of the gfx950 executor (host side).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace atn {

using Leg = uint32_t;
using Path = std::vector<std::pair<size_t, size_t>>;

struct Error : std::runtime_error {
  int code;
  Error(const std::string& s, int c = -1) : std::runtime_error(s), code(c) {}
};

template <class... A>
std::string cat(const A&... a) {
  std::ostringstream o;
  (o << ... << a);
  return o.str();
}

inline std::string vstr(const std::vector<size_t>& v) {
  std::ostringstream o;
  o << "[";
  for (size_t i = 0; i < v.size(); ++i) o << (i ? ", " : "") << v[i];
  o << "]";
  return o.str();
}

#define ATN_HIP(expr)                                                                                \
  do {                                                                                               \
    hipError_t e_ = (expr);                                                                          \
    if (e_ != hipSuccess) throw ::atn::Error(::atn::cat(#expr, " failed: ", hipGetErrorString(e_)), -2); \
  } while (0)

// dtype codes of the C ABI
enum DType { F32 = 0, F64 = 1, C64 = 2, C128 = 3 };
enum Prec { NATIVE = 0, C3M = 1, BF16X3 = 2, FP16X2 = 3, BF16X3_3M = 4, FP16X2_3M = 5 };
// The split precisions with complex products as three real ones (3M) where the
// 3M kernel applies: their half type is that of BF16X3 / FP16X2.
inline bool split_3m(int prec) { return prec == BF16X3_3M || prec == FP16X2_3M; }
inline int split_halves(int prec) { return prec == BF16X3_3M ? BF16X3 : prec == FP16X2_3M ? FP16X2 : prec; }
inline bool split_precision(int prec) { return split_halves(prec) == BF16X3 || split_halves(prec) == FP16X2; }

inline size_t dtype_size(int dt) {
  switch (dt) {
    case F32: return 4;
    case F64: return 8;
    case C64: return 8;
    case C128: return 16;
  }
  throw Error("unsupported dtype code");
}
inline bool dtype_complex(int dt) { return dt == C64 || dt == C128; }

inline size_t checked_numel(const std::vector<size_t>& shape) {
  size_t n = 1;
  for (size_t ax = 0; ax < shape.size(); ++ax) {
    if (shape[ax] == 0) throw Error(cat("shape axis ", ax, " is 0; zero-length axes are not supported"));
    if (n > SIZE_MAX / shape[ax]) throw Error(cat("element count of shape ", vstr(shape), " overflows"));
    n *= shape[ax];
  }
  return n;
}

inline size_t prod(const std::vector<size_t>& v) {
  size_t p = 1;
  for (size_t x : v) p *= x;
  return p;
}

}  // namespace atn
