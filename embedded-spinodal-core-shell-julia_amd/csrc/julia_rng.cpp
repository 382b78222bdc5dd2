// Host-side initial_fields (core_shell.jl:27-37) with Julia's
// rand(MersenneTwister(seed), n, n) reproduced bit for bit.
//
// Julia 1.12's stdlib Random (Manifest.toml:3 pins Julia 1.12.6):
//  * MersenneTwister(seed): SHA-256 over the seed's little-endian 32-bit
//    limbs (least significant first, at least one limb; negative seeds are
//    complemented and get a trailing 0x01 byte); the digest as 8 LE uint32
//    words keys dSFMT's init_by_array.
//  * dSFMT 2.2, Mersenne exponent 19937 (M. Saito, M. Matsumoto).
//  * rand!(::MersenneTwister, ::Array{Float64}) of length n: when
//    n2 = (n-2)/2*2 >= 382, the first n2 values come from one
//    fill_array_close1_open2 call on the generator state and the remaining
//    values from the 1002-value cache, filled afresh from the state left by
//    the bulk call; smaller arrays are served from the cache alone.  Values
//    are [1,2) doubles minus 1.0.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/csg.h"

namespace {

// ---- SHA-256 (FIPS 180-4) --------------------------------------------------
struct Sha256 {
  uint32_t h[8];
  uint8_t buf[64];
  uint64_t len = 0;
  size_t nbuf = 0;

  static uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

  Sha256() {
    static const uint32_t init[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a,
                                     0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
    std::memcpy(h, init, sizeof(h));
  }

  void block(const uint8_t* p) {
    static const uint32_t K[64] = {
        0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
        0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
        0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
        0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
        0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
        0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
        0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
        0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};
    uint32_t w[64];
    for (int i = 0; i < 16; ++i)
      w[i] = (uint32_t(p[4 * i]) << 24) | (uint32_t(p[4 * i + 1]) << 16) | (uint32_t(p[4 * i + 2]) << 8) |
             uint32_t(p[4 * i + 3]);
    for (int i = 16; i < 64; ++i) {
      const uint32_t s0 = rotr(w[i - 15], 7) ^ rotr(w[i - 15], 18) ^ (w[i - 15] >> 3);
      const uint32_t s1 = rotr(w[i - 2], 17) ^ rotr(w[i - 2], 19) ^ (w[i - 2] >> 10);
      w[i] = w[i - 16] + s0 + w[i - 7] + s1;
    }
    uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], hh = h[7];
    for (int i = 0; i < 64; ++i) {
      const uint32_t S1 = rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25);
      const uint32_t ch = (e & f) ^ (~e & g);
      const uint32_t t1 = hh + S1 + ch + K[i] + w[i];
      const uint32_t S0 = rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22);
      const uint32_t mj = (a & b) ^ (a & c) ^ (b & c);
      const uint32_t t2 = S0 + mj;
      hh = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
    }
    h[0] += a; h[1] += b; h[2] += c; h[3] += d; h[4] += e; h[5] += f; h[6] += g; h[7] += hh;
  }

  void update(const uint8_t* p, size_t n) {
    len += n;
    while (n > 0) {
      const size_t k = std::min(n, size_t(64) - nbuf);
      std::memcpy(buf + nbuf, p, k);
      nbuf += k; p += k; n -= k;
      if (nbuf == 64) { block(buf); nbuf = 0; }
    }
  }

  void digest(uint8_t out[32]) {
    const uint64_t bits = len * 8;
    const uint8_t pad = 0x80;
    update(&pad, 1);
    const uint8_t zero = 0;
    while (nbuf != 56) update(&zero, 1);
    uint8_t lb[8];
    for (int i = 0; i < 8; ++i) lb[i] = uint8_t(bits >> (56 - 8 * i));
    update(lb, 8);
    for (int i = 0; i < 8; ++i) {
      out[4 * i] = uint8_t(h[i] >> 24); out[4 * i + 1] = uint8_t(h[i] >> 16);
      out[4 * i + 2] = uint8_t(h[i] >> 8); out[4 * i + 3] = uint8_t(h[i]);
    }
  }
};

// ---- dSFMT-19937 -----------------------------------------------------------
constexpr int MEXP = 19937;
constexpr int NW = (MEXP - 128) / 104 + 1;  // 191 128-bit words
constexpr int N64 = NW * 2;                 // 382
constexpr int POS1 = 117;
constexpr int SL1 = 19;
constexpr int SR = 12;
constexpr uint64_t MSK1 = 0x000ffafffffffb3fULL;
constexpr uint64_t MSK2 = 0x000ffdfffc90fffdULL;
constexpr uint64_t FIX1 = 0x90014964b32f4329ULL;
constexpr uint64_t FIX2 = 0x3b8d12ac548a7c7aULL;
constexpr uint64_t PCV1 = 0x3d84e1ac0dc82880ULL;
constexpr uint64_t PCV2 = 0x0000000000000001ULL;
constexpr uint64_t LOW_MASK = 0x000FFFFFFFFFFFFFULL;
constexpr uint64_t HIGH_CONST = 0x3FF0000000000000ULL;
constexpr int CACHE = 1002;  // Julia MersenneTwister float cache (values)

struct W128 {
  uint64_t u[2];
};

inline void recursion(W128* r, const W128* a, const W128* b, W128* lung) {
  const uint64_t t0 = a->u[0], t1 = a->u[1];
  const uint64_t L0 = lung->u[0], L1 = lung->u[1];
  lung->u[0] = (t0 << SL1) ^ (L1 >> 32) ^ (L1 << 32) ^ b->u[0];
  lung->u[1] = (t1 << SL1) ^ (L0 >> 32) ^ (L0 << 32) ^ b->u[1];
  r->u[0] = (lung->u[0] >> SR) ^ (lung->u[0] & MSK1) ^ t0;
  r->u[1] = (lung->u[1] >> SR) ^ (lung->u[1] & MSK2) ^ t1;
}

inline uint32_t ini1(uint32_t x) { return (x ^ (x >> 27)) * 1664525u; }
inline uint32_t ini2(uint32_t x) { return (x ^ (x >> 27)) * 1566083941u; }

struct Dsfmt {
  W128 st[NW + 1];

  void init_by_array(const uint32_t* key, int key_length) {
    const int size = (NW + 1) * 4;
    const int lag = 11;
    const int mid = (size - lag) / 2;
    uint32_t s32[(NW + 1) * 4];
    for (int i = 0; i < size; ++i) s32[i] = 0x8b8b8b8bu;
    int count = (key_length + 1 > size) ? key_length + 1 : size;
    uint32_t r = ini1(s32[0] ^ s32[mid % size] ^ s32[(size - 1) % size]);
    s32[mid % size] += r;
    r += static_cast<uint32_t>(key_length);
    s32[(mid + lag) % size] += r;
    s32[0] = r;
    count--;
    int i = 1, j = 0;
    for (; j < count && j < key_length; ++j) {
      r = ini1(s32[i] ^ s32[(i + mid) % size] ^ s32[(i + size - 1) % size]);
      s32[(i + mid) % size] += r;
      r += key[j] + static_cast<uint32_t>(i);
      s32[(i + mid + lag) % size] += r;
      s32[i] = r;
      i = (i + 1) % size;
    }
    for (; j < count; ++j) {
      r = ini1(s32[i] ^ s32[(i + mid) % size] ^ s32[(i + size - 1) % size]);
      s32[(i + mid) % size] += r;
      r += static_cast<uint32_t>(i);
      s32[(i + mid + lag) % size] += r;
      s32[i] = r;
      i = (i + 1) % size;
    }
    for (j = 0; j < size; ++j) {
      r = ini2(s32[i] + s32[(i + mid) % size] + s32[(i + size - 1) % size]);
      s32[(i + mid) % size] ^= r;
      r -= static_cast<uint32_t>(i);
      s32[(i + mid + lag) % size] ^= r;
      s32[i] = r;
      i = (i + 1) % size;
    }
    for (int k = 0; k <= NW; ++k) {
      st[k].u[0] = uint64_t(s32[4 * k]) | (uint64_t(s32[4 * k + 1]) << 32);
      st[k].u[1] = uint64_t(s32[4 * k + 2]) | (uint64_t(s32[4 * k + 3]) << 32);
    }
    for (int k = 0; k < NW; ++k)
      for (int q = 0; q < 2; ++q) st[k].u[q] = (st[k].u[q] & LOW_MASK) | HIGH_CONST;
    // period certification
    const uint64_t t0 = st[NW].u[0] ^ FIX1, t1 = st[NW].u[1] ^ FIX2;
    uint64_t inner = (t0 & PCV1) ^ (t1 & PCV2);
    for (int s = 32; s > 0; s >>= 1) inner ^= inner >> s;
    if (!(inner & 1)) st[NW].u[1] ^= 1;
  }

  // gen_rand_array_c1o2: `size` 128-bit words into arr, advancing the state.
  void fill(W128* arr, int size) {
    W128 lung = st[NW];
    int i, j;
    recursion(&arr[0], &st[0], &st[POS1], &lung);
    for (i = 1; i < NW - POS1; ++i) recursion(&arr[i], &st[i], &st[i + POS1], &lung);
    for (; i < NW; ++i) recursion(&arr[i], &st[i], &arr[i + POS1 - NW], &lung);
    for (; i < size - NW; ++i) recursion(&arr[i], &arr[i - NW], &arr[i + POS1 - NW], &lung);
    for (j = 0; j < 2 * NW - size; ++j) st[j] = arr[j + size - NW];
    for (; i < size; ++i, ++j) {
      recursion(&arr[i], &arr[i - NW], &arr[i + POS1 - NW], &lung);
      st[j] = arr[i];
    }
    st[NW] = lung;
  }
};

struct JuliaMT {
  Dsfmt d;
  std::vector<W128> cache;
  int idx = CACHE;  // empty cache

  explicit JuliaMT(int64_t seed) {
    Sha256 sha;
    bool neg = seed < 0;
    uint64_t s = neg ? static_cast<uint64_t>(~seed) : static_cast<uint64_t>(seed);
    while (true) {
      const uint32_t w = static_cast<uint32_t>(s & 0xffffffffu);
      const uint8_t b[4] = {uint8_t(w), uint8_t(w >> 8), uint8_t(w >> 16), uint8_t(w >> 24)};
      sha.update(b, 4);
      s >>= 32;
      if (s == 0) break;
    }
    if (neg) {
      const uint8_t one = 1;
      sha.update(&one, 1);
    }
    uint8_t dg[32];
    sha.digest(dg);
    uint32_t key[8];
    for (int i = 0; i < 8; ++i)
      key[i] = uint32_t(dg[4 * i]) | (uint32_t(dg[4 * i + 1]) << 8) | (uint32_t(dg[4 * i + 2]) << 16) |
               (uint32_t(dg[4 * i + 3]) << 24);
    d.init_by_array(key, 8);
    cache.resize(CACHE / 2);
  }

  double scalar() {
    if (idx >= CACHE) {
      d.fill(cache.data(), CACHE / 2);
      idx = 0;
    }
    uint64_t bits = cache[idx / 2].u[idx % 2];
    ++idx;
    double v;
    std::memcpy(&v, &bits, 8);
    return v - 1.0;
  }

  void rand(double* out, int64_t n) {
    const int64_t n2 = (n - 2) / 2 * 2;
    if (n2 < N64) {
      for (int64_t k = 0; k < n; ++k) out[k] = scalar();
      return;
    }
    std::vector<W128> bulk(static_cast<size_t>(n2 / 2));
    d.fill(bulk.data(), static_cast<int>(n2 / 2));
    for (int64_t k = 0; k < n2; ++k) {
      double v;
      std::memcpy(&v, &bulk[static_cast<size_t>(k / 2)].u[k % 2], 8);
      out[k] = v - 1.0;
    }
    for (int64_t k = n2; k < n; ++k) out[k] = scalar();
  }
};

// Julia's sum over a Float64 vector: pairwise halving above 1024 values,
// sequential below (Julia vectorises that part; last-ulp differences).
double pairwise_sum(const double* x, int64_t n) {
  if (n <= 1024) {
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += x[i];
    return s;
  }
  const int64_t h = ((n - 1) >> 1) + 1;
  return pairwise_sum(x, h) + pairwise_sum(x + h, n - h);
}

thread_local std::string g_rng_err;

}  // namespace

extern "C" int csg_julia_rand(int64_t seed, int64_t count, double* out) {
  if (count < 0 || (count > 0 && !out)) return CSG_ERR_ARG;
  if (count == 0) return CSG_OK;
  if (count > (int64_t(1) << 31) * 2) return CSG_ERR_ARG;
  JuliaMT mt(seed);
  mt.rand(out, count);
  return CSG_OK;
}

extern "C" int csg_initial_fields(int32_t n, double spacing, double radius, double width, int64_t seed,
                                  double* c_out, double* phi_out) {
  if (n < 1 || !c_out || !phi_out || !(width != 0.0)) return CSG_ERR_ARG;
  const size_t nn = static_cast<size_t>(n) * n;
  // phi[i, j] = 0.5(1 - tanh((hypot((i-1-n/2)dx, (j-1-n/2)dx) - R)/w)), 1-based i, j
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      const double x = static_cast<double>(i - n / 2) * spacing;
      const double y = static_cast<double>(j - n / 2) * spacing;
      phi_out[i + static_cast<size_t>(j) * n] = 0.5 * (1.0 - std::tanh((std::hypot(x, y) - radius) / width));
    }
  std::vector<double> noise(nn);
  JuliaMT mt(seed);
  mt.rand(noise.data(), static_cast<int64_t>(nn));
  for (size_t k = 0; k < nn; ++k) noise[k] = 0.01 * (2.0 * noise[k] - 1.0);
  // mean(noise[inside]) over the column-major gather
  std::vector<double> sel;
  sel.reserve(nn);
  for (size_t k = 0; k < nn; ++k)
    if (phi_out[k] >= 0.5) sel.push_back(noise[k]);
  const double m = sel.empty() ? NAN : pairwise_sum(sel.data(), static_cast<int64_t>(sel.size())) / sel.size();
  for (size_t k = 0; k < nn; ++k) {
    c_out[k] = 0.5;
    if (phi_out[k] >= 0.5) c_out[k] += noise[k] - m;
  }
  return CSG_OK;
}
