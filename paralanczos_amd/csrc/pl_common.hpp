// Shared host/device definitions for the batched-submatrix Lanczos path:
// status bits, the breakdown threshold, and the counter-based PRNG streams
// from which subset index sets and synthetic sensing matrices are drawn.
//
// Stream format (identical on host and device, bit-exact):
//   word w of stream (seed, tag, sid) = Philox4x32-10 with
//     counter = {w / 4, sid & 0xffffffff, sid >> 32, tag},
//     key     = {seed & 0xffffffff, seed >> 32},
//   lane w % 4 of the result.
//   Philox4x32-10: Salmon, Moraes, Dror, Shaw, SC'11.
//   Bounded draws use Lemire's multiply-shift with rejection (ACM TOMACS 2019);
//   index sets use Floyd's sampling algorithm (CACM 1987), kept in draw order.
// This replaces the reference's curand/clock() start vectors
// (ParaLanczos.cu:1196-1206); the reference takes caller-supplied index sets
// (ParaLanczos.cu:1811, example3.cu:43-49) and has no generator of its own.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define PL_HD __host__ __device__ inline
#else
#define PL_HD inline
#endif

namespace pl {

// per-subset status word
constexpr uint32_t ST_BREAKDOWN = 1u;  // Lanczos restart after beta <= tau * scale
constexpr uint32_t ST_NONFINITE = 2u;  // non-finite value in G_S or in T
constexpr uint32_t ST_NOCONV = 4u;     // bisection reached its iteration cap
constexpr uint32_t ST_BADINDEX = 8u;   // index outside [0, n)
constexpr uint32_t ST_BADK = 16u;      // subset size outside the supported range

// restart when beta_{j+1} <= TAU * (max_i |alpha_i| + max_i beta_i), i.e. only
// for an (almost) exactly vanishing beta. A small nonzero beta is kept: the
// full reorthogonalisation keeps the noise-led next direction orthogonal.
constexpr float TAU = 1.0f / 1073741824.0f;

enum : uint32_t { TAG_IDX = 1u, TAG_K = 2u, TAG_PHI = 3u, TAG_DCT = 4u };

PL_HD void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int round = 0; round < 10; ++round) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    c[1] = (uint32_t)p1;
    c[3] = (uint32_t)p0;
    c[0] = n0;
    c[2] = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// one 128-bit block of stream (seed, tag, sid)
PL_HD void stream_block(uint64_t seed, uint32_t tag, uint64_t sid, uint32_t blk,
                        uint32_t out[4]) {
  out[0] = blk;
  out[1] = (uint32_t)sid;
  out[2] = (uint32_t)(sid >> 32);
  out[3] = tag;
  philox4x32_10(out, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// sequential reader of one stream
struct Stream {
  uint64_t seed, sid;
  uint32_t tag, w;
  uint32_t buf[4];

  PL_HD Stream(uint64_t seed_, uint32_t tag_, uint64_t sid_)
      : seed(seed_), sid(sid_), tag(tag_), w(0u), buf{0u, 0u, 0u, 0u} {}

  PL_HD uint32_t next() {
    if ((w & 3u) == 0u) stream_block(seed, tag, sid, w >> 2, buf);
    return buf[(w++) & 3u];
  }

  // uniform in [0, range), range >= 1
  PL_HD uint32_t bounded(uint32_t range) {
    uint64_t m = (uint64_t)next() * range;
    uint32_t lo = (uint32_t)m;
    if (lo < range) {
      const uint32_t thresh = (0u - range) % range;
      while (lo < thresh) {
        m = (uint64_t)next() * range;
        lo = (uint32_t)m;
      }
    }
    return (uint32_t)(m >> 32);
  }
};

// 53-bit uniform double in [0, 1) from two words
PL_HD double uniform53(uint32_t hi, uint32_t lo) {
  return ((double)(hi >> 5) * 67108864.0 + (double)(lo >> 6)) * (1.0 / 9007199254740992.0);
}

}  // namespace pl
