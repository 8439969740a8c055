// Random-number parity with the reference's host generators, and the probit transform of the
// auxiliary fields on the GPU.
//
//   split_seed (utilities/Random.cpp:7-20): the (seed, rank, stream) triple packed as
//     seed << 32 | stream << 24 | rank and passed through the splitmix64 finalizer
//     (Steele, Lea, Flood 2014; constants as published with splitmix64).
//   HostRandomGenerator (utilities/Random.hpp:36-49): std::mt19937 seeded with that 64-bit
//     value (the engine keeps it modulo 2^32), uniforms from
//     std::uniform_real_distribution<double>(0, 1). The engine is the 32-bit Mersenne Twister
//     of Matsumoto and Nishimura (1998) with the initialisation of the C++ standard
//     ([rand.eng.mers]); a uniform is libstdc++'s generate_canonical<double, 53> over two draws
//     x1, x2: (x1 + x2 2^32) / 2^64 in double arithmetic, and the largest double below 1 if
//     that rounds to 1.
//   probit (AFQMC/Utilities/probit.h:57-135): Acklam's rational approximation of the inverse
//     normal CDF (P. J. Acklam, "An algorithm for computing the inverse normal cumulative
//     distribution function", 2002; relative error below 1.15e-9), applied on the device to the
//     uniforms of a step before construct_X (construct_X_kernel.hpp:71).
//
// The generators run on the host, as the reference's do (FromHostRandomGenerator for device
// builds, Random.hpp:87-99): the draws are sequential and a step needs only W*K of them.
#include "afq_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <new>
#include <string>

struct afq_rng_t {
  uint32_t mt[624];
  int idx;
};

namespace afq {

namespace {

constexpr int kN = 624, kM = 397;
constexpr uint32_t kMatrixA = 0x9908b0dfu, kUpper = 0x80000000u, kLower = 0x7fffffffu;

void mt_seed(afq_rng_t* r, uint64_t seed) {
  r->mt[0] = (uint32_t)(seed & 0xffffffffu);
  for (int i = 1; i < kN; ++i)
    r->mt[i] = 1812433253u * (r->mt[i - 1] ^ (r->mt[i - 1] >> 30)) + (uint32_t)i;
  r->idx = kN;
}

void mt_twist(afq_rng_t* r) {
  for (int i = 0; i < kN; ++i) {
    const uint32_t y = (r->mt[i] & kUpper) | (r->mt[(i + 1) % kN] & kLower);
    r->mt[i] = r->mt[(i + kM) % kN] ^ (y >> 1) ^ ((y & 1u) ? kMatrixA : 0u);
  }
  r->idx = 0;
}

uint32_t mt_next(afq_rng_t* r) {
  if (r->idx >= kN) mt_twist(r);
  uint32_t y = r->mt[r->idx++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

double mt_uniform(afq_rng_t* r) {
  double sum = (double)mt_next(r);
  sum += (double)mt_next(r) * 4294967296.0;
  double u = sum / 18446744073709551616.0;
  if (u >= 1.0) u = std::nextafter(1.0, 0.0);
  return u;
}

__device__ inline double probit(double p) {
  // evaluated as written, each product and sum rounded on its own (no fused multiply-add), as
  // a plain reading of the source computes it
#pragma clang fp contract(off)
  const double a0 = -3.969683028665376e+01, a1 = 2.209460984245205e+02,
               a2 = -2.759285104469687e+02, a3 = 1.383577518672690e+02,
               a4 = -3.066479806614716e+01, a5 = 2.506628277459239e+00;
  const double b0 = -5.447609879822406e+01, b1 = 1.615858368580409e+02,
               b2 = -1.556989798598866e+02, b3 = 6.680131188771972e+01,
               b4 = -1.328068155288572e+01;
  const double c0 = -7.784894002430293e-03, c1 = -3.223964580411365e-01,
               c2 = -2.400758277161838e+00, c3 = -2.549732539343734e+00,
               c4 = 4.374664141464968e+00, c5 = 2.938163982698783e+00;
  const double d0 = 7.784695709041462e-03, d1 = 3.224671290700398e-01,
               d2 = 2.445134137142996e+00, d3 = 3.754408661907416e+00;
  const double lo = 0.02425, hi = 0.97575;   // the reference's PROBIT_LOW, PROBIT_HIGH
  if (p < 0.0 || p > 1.0) return 0.0;
  if (p == 0.0) return -HUGE_VAL;
  if (p == 1.0) return HUGE_VAL;
  if (p < lo) {
    const double q = sqrt(-2.0 * log(p));
    return (((((c0 * q + c1) * q + c2) * q + c3) * q + c4) * q + c5) /
           ((((d0 * q + d1) * q + d2) * q + d3) * q + 1.0);
  }
  if (p > hi) {
    const double q = sqrt(-2.0 * log(1.0 - p));
    return -(((((c0 * q + c1) * q + c2) * q + c3) * q + c4) * q + c5) /
           ((((d0 * q + d1) * q + d2) * q + d3) * q + 1.0);
  }
  const double q = p - 0.5, r = q * q;
  return (((((a0 * r + a1) * r + a2) * r + a3) * r + a4) * r + a5) * q /
         (((((b0 * r + b1) * r + b2) * r + b3) * r + b4) * r + 1.0);
}

int fail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}

__global__ __launch_bounds__(256) void k_probit(long n, const double* __restrict__ u,
                                                double* __restrict__ x) {
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256)
    x[i] = probit(u[i]);
}

}  // namespace

}  // namespace afq

using namespace afq;

extern "C" {

unsigned long long afq_split_seed(int seed, int rank, unsigned stream) {
  uint64_t x = ((uint64_t)(unsigned)seed << 32) | ((uint64_t)(stream & 0xffu) << 24) |
               (uint64_t)(unsigned)rank;
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

int afq_rng_create(unsigned long long seed, afq_rng* out) {
  if (out == nullptr) return fail(AFQ_EINVAL, "afq_rng_create: null output handle");
  afq_rng_t* r = new (std::nothrow) afq_rng_t;
  if (r == nullptr) return fail(AFQ_ENOMEM, "afq_rng_create: out of host memory");
  mt_seed(r, seed);
  *out = r;
  return AFQ_OK;
}

void afq_rng_destroy(afq_rng r) { delete r; }

int afq_rng_raw(afq_rng r, long n, unsigned* out) {
  if (r == nullptr || n < 0 || (n > 0 && out == nullptr))
    return fail(AFQ_EINVAL, "afq_rng_raw: null generator or output, or negative count");
  for (long i = 0; i < n; ++i) out[i] = mt_next(r);
  return AFQ_OK;
}

int afq_rng_uniform(afq_rng r, long n, double* out) {
  if (r == nullptr || n < 0 || (n > 0 && out == nullptr))
    return fail(AFQ_EINVAL, "afq_rng_uniform: null generator or output, or negative count");
  for (long i = 0; i < n; ++i) out[i] = mt_uniform(r);
  return AFQ_OK;
}

int afq_probit(long n, const double* d_u, double* d_x, void* stream) {
  if (n < 0 || (n > 0 && (d_u == nullptr || d_x == nullptr)))
    return fail(AFQ_EINVAL, "afq_probit: null operand or negative count");
  if (n == 0) return AFQ_OK;
  const long blocks = std::min<long>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_probit, dim3((unsigned)blocks), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), n, d_u, d_x);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(AFQ_EHIP, std::string("afq_probit: ") + hipGetErrorString(e));
  return AFQ_OK;
}

}  // extern "C"
