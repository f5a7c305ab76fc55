// SplitMix64 stream and the MurmurHash3 64-bit finaliser: the only sources of
// pseudo-randomness and hashing in the library, so generated inputs and set
// hashes are identical on every x86-64 host.
#pragma once
#include <cstdint>

namespace gpsm {

inline uint64_t fmix64(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdULL;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ULL;
  k ^= k >> 33;
  return k;
}

// Per-vertex term of the tuple hash: a bijection of (query id, data id).
inline uint64_t tuple_term(uint32_t qid, uint32_t v) {
  return fmix64((((uint64_t)qid << 32) | v) ^ 0x9E3779B97F4A7C15ULL) | 1ULL;
}

struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  SplitMix64(uint64_t seed, uint64_t stream) : s(fmix64(seed) ^ fmix64(stream + 0x632BE59BD9B4E019ULL)) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  // uniform in [0, n) by multiply-shift (Lemire); n > 0
  uint64_t below(uint64_t n) { return (uint64_t)(((unsigned __int128)next() * n) >> 64); }
  // uniform double in [0, 1) with 53 random bits
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

}  // namespace gpsm
