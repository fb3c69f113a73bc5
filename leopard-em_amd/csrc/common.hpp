// Host-side helpers shared by the translation units of libtdtm: the thread-local error
// message behind tdtm_last_error(), HIP status checks, and the twiddle table.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "fft.hpp"

namespace tdtm {

// one message per host thread, shared by every translation unit
inline std::string& last_error() {
  thread_local std::string s;
  return s;
}

inline int fail(const std::string& msg) {
  last_error() = msg;
  return 1;
}

#define TDTM_HIP(call)                                                                  \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) return fail(std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define TDTM_LAUNCHED()                                                                 \
  do {                                                                                  \
    hipError_t e_ = hipGetLastError();                                                  \
    if (e_ != hipSuccess) return fail(std::string("kernel launch: ") + hipGetErrorString(e_)); \
  } while (0)

inline bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

// tw[m] = e^{-2 pi i m / TW_N}: evaluated on the first octant and mapped by symmetry,
// so quarter and eighth turns are exact and the table has the circle's symmetries
inline std::vector<float2> twiddle_table() {
  std::vector<float2> tw(TW_N);
  constexpr int Q = TW_N / 4;
  for (int m = 0; m < TW_N; ++m) {
    const int q = m / Q, r = m % Q;
    double c, sn;  // cos, sin of 2 pi r / TW_N, r in [0, Q)
    if (2 * r <= Q) {
      const double a = 2.0 * M_PI * (double)r / (double)TW_N;
      c = std::cos(a);
      sn = std::sin(a);
    } else {
      const double a = 2.0 * M_PI * (double)(Q - r) / (double)TW_N;
      c = std::sin(a);
      sn = std::cos(a);
    }
    double cq = c, sq = sn;  // rotate by q quarter turns
    if (q == 1) cq = -sn, sq = c;
    if (q == 2) cq = -c, sq = -sn;
    if (q == 3) cq = sn, sq = -c;
    tw[m] = make_float2((float)cq, (float)-sq);
  }
  return tw;
}

}  // namespace tdtm
