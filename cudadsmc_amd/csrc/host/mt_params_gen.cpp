// Generator of the Mersenne-Twister parameter table used by the DSMC RNG
// streams (writes mt_params_table.inc next to this file).
//
// The generator family is the one the reference draws its per-stream
// parameters from (kernels/mtwister.cu:70-100, include/MersenneTwister.h:
// 17-40): word size w = 32, degree n = 19, middle m = 9, r = 1 separation bit,
// so the state holds p = n*w - r = 607 bits and the period is 2^607 - 1 (a
// Mersenne prime). Parameter sets are searched the way Matsumoto & Nishimura
// describe for dynamic creation (DCMT, 1998): draw a random twist vector
// `a` (top bit set so the twist matrix is invertible), obtain the
// characteristic polynomial of the state recurrence from 2p output bits with
// the Berlekamp-Massey algorithm, reject candidates with a small factor, and
// accept when x^(2^p) == x mod phi(x). Because p and 2^p - 1 are prime this
// proves phi irreducible, hence primitive, hence the full period.
//
// Tempering uses the shifts of the reference family (12, 7, 15, 18) and the
// published MT19937 tempering masks b = 0x9d2c5680, c = 0xefc60000 for every
// set (DCMT would additionally optimise b and c per set for
// equidistribution; that refinement is not done here).
//
// Build and run:  g++ -O2 -std=c++17 -pthread mt_params_gen.cpp -o gen && ./gen 1024
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>
#include <atomic>
#include <algorithm>

namespace {

constexpr int P = 607;          // state bits
constexpr int NN = 19, MM = 9;  // words, middle offset
constexpr int WORDS = 20;       // 1280 bits: room for products of degree < 1214

struct Poly {
  uint64_t w[WORDS];
  void clear() { std::memset(w, 0, sizeof(w)); }
  int bit(int i) const { return (int)((w[i >> 6] >> (i & 63)) & 1u); }
  void set(int i) { w[i >> 6] |= (uint64_t)1 << (i & 63); }
  void flip(int i) { w[i >> 6] ^= (uint64_t)1 << (i & 63); }
  int degree() const {
    for (int k = WORDS - 1; k >= 0; --k)
      if (w[k]) return k * 64 + 63 - __builtin_clzll(w[k]);
    return -1;
  }
};

// r ^= a << s
void xor_shifted(Poly& r, const Poly& a, int s) {
  int ws = s >> 6, bs = s & 63;
  for (int k = WORDS - 1; k >= ws; --k) {
    uint64_t v = a.w[k - ws] << bs;
    if (bs && k - ws - 1 >= 0) v |= a.w[k - ws - 1] >> (64 - bs);
    r.w[k] ^= v;
  }
}

uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Berlekamp-Massey over GF(2); returns the connection polynomial C(x) and L.
int berlekamp_massey(const std::vector<uint8_t>& s, Poly& C) {
  Poly B, T, W;
  C.clear(); B.clear(); W.clear();
  C.set(0); B.set(0);
  int L = 0, m = 1;
  const int N = (int)s.size();
  for (int n = 0; n < N; ++n) {
    // W[i] = s[n - i]
    for (int k = WORDS - 1; k > 0; --k) W.w[k] = (W.w[k] << 1) | (W.w[k - 1] >> 63);
    W.w[0] = (W.w[0] << 1) | s[n];
    int d = 0;
    for (int k = 0; k < WORDS; ++k) d ^= __builtin_popcountll(C.w[k] & W.w[k]) & 1;
    if (!d) { ++m; continue; }
    if (2 * L <= n) {
      T = C;
      xor_shifted(C, B, m);
      L = n + 1 - L;
      B = T;
      m = 1;
    } else {
      xor_shifted(C, B, m);
      ++m;
    }
  }
  return L;
}

// remainder of f modulo a small polynomial q of degree dq (< 32)
uint32_t mod_small(const Poly& f, int deg_f, uint32_t q, int dq) {
  uint32_t r = 0;
  for (int i = deg_f; i >= 0; --i) {
    r = (r << 1) | (uint32_t)f.bit(i);
    if (r >> dq & 1u) r ^= q;
  }
  return r;
}

std::vector<std::pair<uint32_t, int>> small_irreducibles(int maxdeg) {
  std::vector<std::pair<uint32_t, int>> out;
  for (int d = 1; d <= maxdeg; ++d) {
    for (uint32_t q = (1u << d); q < (2u << d); ++q) {
      bool irr = true;
      for (auto& pr : out) {
        // polynomial division remainder of q by pr
        uint32_t r = 0;
        for (int i = d; i >= 0; --i) {
          r = (r << 1) | ((q >> i) & 1u);
          if (r >> pr.second & 1u) r ^= pr.first;
        }
        if (r == 0) { irr = false; break; }
      }
      if (irr) out.push_back({q, d});
    }
  }
  return out;
}

// r = r^2 mod f (deg f == P)
void square_mod(Poly& r, const Poly& f) {
  Poly sq;
  sq.clear();
  for (int k = 0; k < WORDS / 2; ++k) {
    uint64_t v = r.w[k];
    uint64_t lo = 0, hi = 0;
    for (int b = 0; b < 32; ++b) {
      lo |= ((v >> b) & 1ull) << (2 * b);
      hi |= ((v >> (b + 32)) & 1ull) << (2 * b);
    }
    sq.w[2 * k] = lo;
    sq.w[2 * k + 1] = hi;
  }
  for (int j = 2 * (P - 1); j >= P; --j)
    if (sq.bit(j)) xor_shifted(sq, f, j - P);
  r = sq;
}

bool is_irreducible_deg_p(const Poly& f) {
  Poly r;
  r.clear();
  r.set(1);
  for (int i = 0; i < P; ++i) square_mod(r, f);
  Poly x;
  x.clear();
  x.set(1);
  return std::memcmp(r.w, x.w, sizeof(r.w)) == 0;
}

uint32_t twist(uint32_t lo_word, uint32_t next_word, uint32_t a) {
  uint32_t x = (lo_word & 0xFFFFFFFEu) | (next_word & 1u);
  return (x >> 1) ^ ((x & 1u) ? a : 0u);
}

bool candidate_ok(uint32_t a, uint64_t& rng,
                  const std::vector<std::pair<uint32_t, int>>& smalls) {
  uint32_t st[NN];
  for (int i = 0; i < NN; ++i) st[i] = (uint32_t)splitmix(rng);
  std::vector<uint8_t> bits(2 * P);
  int idx = 0;  // block form of the recurrence
  for (int n = 0; n < 2 * P; ++n) {
    int k = idx, k1 = (idx + 1) % NN, km = (idx + MM) % NN;
    uint32_t x = st[km] ^ twist(st[k], st[k1], a);
    st[k] = x;
    idx = k1;
    bits[n] = (uint8_t)(x >> 31);
  }
  Poly C;
  int L = berlekamp_massey(bits, C);
  if (L != P || C.degree() != P) return false;
  for (auto& q : smalls)
    if (mod_small(C, P, q.first, q.second) == 0) return false;
  return is_irreducible_deg_p(C);
}

}  // namespace

int main(int argc, char** argv) {
  if (argc > 1 && std::strcmp(argv[1], "--check") == 0) {
    // verify given twist vectors: prints "<vector> 1" when the polynomial is primitive
    auto smalls = small_irreducibles(14);
    uint64_t rng = 0x1234567ull;
    for (int i = 2; i < argc; ++i) {
      uint32_t a = (uint32_t)std::strtoul(argv[i], nullptr, 16);
      std::printf("%s %d\n", argv[i], candidate_ok(a, rng, smalls) ? 1 : 0);
    }
    return 0;
  }
  int count = argc > 1 ? std::atoi(argv[1]) : 1024;
  const char* out_path = argc > 2 ? argv[2] : "mt_params_table.inc";
  auto smalls = small_irreducibles(14);
  std::vector<uint32_t> table(count, 0);
  std::atomic<int> next{0};
  unsigned nth = std::max(1u, std::thread::hardware_concurrency());
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < nth; ++t) {
    pool.emplace_back([&]() {
      for (int i = next++; i < count; i = next++) {
        uint64_t rng = 0x5DEECE66Dull * (uint64_t)(i + 1) + 0x2545F4914F6CDD1Dull;
        for (;;) {
          uint32_t a = (uint32_t)splitmix(rng) | 0x80000000u;
          // embed the table index in the low 16 bits so every row differs
          a = (a & 0xFFFF0000u) | (uint32_t)(i & 0xFFFF);
          if (candidate_ok(a, rng, smalls)) { table[i] = a; break; }
        }
      }
    });
  }
  for (auto& th : pool) th.join();
  FILE* f = std::fopen(out_path, "w");
  if (!f) return 1;
  std::fprintf(f, "// Generated by mt_params_gen.cpp: %d MT(607; n=19, m=9) twist vectors,\n", count);
  std::fprintf(f, "// each with a primitive characteristic polynomial.\n");
  for (int i = 0; i < count; ++i) std::fprintf(f, "0x%08Xu,%s", table[i], (i % 6 == 5) ? "\n" : " ");
  std::fprintf(f, "\n");
  std::fclose(f);
  return 0;
}
