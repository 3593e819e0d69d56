// mt_jump.h -- jumping std::mt19937 ahead (host side, no HIP).
//
// MT19937 is a linear recurrence over GF(2): in stream form its words satisfy
// x[k] = x[k-227] ^ twist(x[k-624], x[k-623]), and the window
// W_n = (x[n] (upper bit), x[n+1], ..., x[n+623]) is a 19937-bit state whose
// transition F has a characteristic polynomial phi of degree 19937.  Every
// linear function of the state therefore satisfies the recurrence phi, and
// with g = t^J mod phi,
//
//     x[n + J + j] = XOR over i with g_i = 1 of x[n + i + j]      (j >= 1)
//
// (for j = 0 the upper bit only, the only part of x[n + J] the window uses).
// So the window J words ahead is a GF(2) correlation of g with the next
// 19937 + 623 words of the stream.  This is the jump-ahead method of
// Haramoto, Matsumoto, Nishimura, Panneton and L'Ecuyer, "Efficient jump
// ahead for F2-linear random number generators" (INFORMS J. Computing 20(3),
// 2008), in the sliding-window form.
//
// phi is found here, once per process, by the Berlekamp-Massey algorithm over
// GF(2) on 2 x 19937 bits of one MT19937 stream (phi is irreducible, so the
// minimal polynomial of any nonzero output bit sequence is phi itself).  The
// jump polynomials t^(2^k) mod phi follow by repeated squaring.
#pragma once
#include <stdint.h>

#include <mutex>
#include <vector>

namespace mtj {

constexpr int kN = 624;         // state words
constexpr int kStep = 227;      // x[k] depends on x[k-227], x[k-624], x[k-623]
constexpr int kDeg = 19937;     // degree of phi
constexpr int kPolyWords = (kDeg + 63) / 64;  // bits 0 .. kDeg-1
constexpr int kStreamWords = kDeg - 1 + kN;   // x[n .. n + 19936 + 623]: one jump's input
constexpr int kMaxPow2 = 255;                  // jumps of 2^k words, k <= kMaxPow2

using Poly = std::vector<uint64_t>;

inline uint32_t twist(uint32_t a, uint32_t b) {
    const uint32_t y = (a & 0x80000000u) | (b & 0x7fffffffu);
    return (y >> 1) ^ ((b & 1u) ? 0x9908b0dfu : 0u);
}

inline uint32_t temper(uint32_t y) {
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

// x[0 .. n) with x[0 .. 624) = w: the raw (untempered) stream from window w.
inline void stream_from(const uint32_t* w, uint32_t* x, int64_t n) {
    for (int64_t k = 0; k < n; ++k)
        x[k] = k < kN ? w[k] : x[k - kStep] ^ twist(x[k - kN], x[k - kN + 1]);
}

// std::mt19937's seeding (the state a fresh generator draws from next).
inline void init_genrand(uint32_t seed, uint32_t* w) {
    w[0] = seed;
    for (int i = 1; i < kN; ++i) w[i] = 1812433253u * (w[i - 1] ^ (w[i - 1] >> 30)) + (uint32_t)i;
}

inline int bit(const Poly& p, int64_t i) { return (int)((p[(size_t)(i >> 6)] >> (i & 63)) & 1u); }

// dst ^= src * t^sh, over the words dst has
inline void xor_shifted(Poly& dst, const Poly& src, int64_t sh) {
    const int64_t ws = sh >> 6;
    const int bs = (int)(sh & 63);
    for (size_t w = 0; w < src.size(); ++w) {
        const uint64_t v = src[w];
        if (!v) continue;
        const size_t a = w + (size_t)ws;
        if (a < dst.size()) dst[a] ^= v << bs;
        if (bs && a + 1 < dst.size()) dst[a + 1] ^= v >> (64 - bs);
    }
}

// Berlekamp-Massey over GF(2): the shortest C (C_0 = 1, degree <= L) with
// sum_{i=0..L} C_i s[n-i] = 0 for all L <= n < s.size(); returns C, sets L.
inline Poly berlekamp_massey(const std::vector<uint8_t>& s, int& L) {
    const int64_t M = (int64_t)s.size();
    const size_t words = (size_t)((M + 64) / 64) + 1;
    // R holds the sequence reversed: bit q is s[M-1-q], so the bits
    // s[n], s[n-1], ..., s[n-L] are R's bits from q0 = M-1-n upwards
    Poly R(words, 0);
    for (int64_t q = 0; q < M; ++q)
        if (s[(size_t)(M - 1 - q)]) R[(size_t)(q >> 6)] |= 1ull << (q & 63);
    const auto window = [&](int64_t q0, size_t w) {  // 64 bits of R from bit q0 + 64w
        const int64_t q = q0 + 64 * (int64_t)w;
        const size_t a = (size_t)(q >> 6);
        const int b = (int)(q & 63);
        const uint64_t lo = a < words ? R[a] : 0;
        const uint64_t hi = a + 1 < words ? R[a + 1] : 0;
        return b ? (lo >> b) | (hi << (64 - b)) : lo;
    };
    Poly C(words, 0), B(words, 0), T;
    C[0] = 1;
    B[0] = 1;
    L = 0;
    int64_t m = -1;
    for (int64_t n = 0; n < M; ++n) {
        const size_t used = (size_t)(L / 64) + 1;
        uint64_t acc = 0;
        for (size_t w = 0; w < used; ++w) acc ^= C[w] & window(M - 1 - n, w);
        if (!(__builtin_popcountll(acc) & 1)) continue;
        if (2 * L <= n) {
            T = C;
            xor_shifted(C, B, n - m);
            L = (int)(n + 1 - L);
            B = T;
            m = n;
        } else {
            xor_shifted(C, B, n - m);
        }
    }
    return C;
}

// The characteristic polynomial phi (bits 0 .. kDeg; bit kDeg set).
inline Poly characteristic_polynomial() {
    std::vector<uint32_t> x((size_t)(2 * kDeg + kN + 2));
    uint32_t w[kN];
    init_genrand(5489u, w);
    stream_from(w, x.data(), (int64_t)x.size());
    // s[n] = top bit of x[n+1], a linear function of the window W_n
    std::vector<uint8_t> s((size_t)(2 * kDeg));
    for (size_t n = 0; n < s.size(); ++n) s[n] = (uint8_t)(x[n + 1] >> 31);
    int L = 0;
    const Poly C = berlekamp_massey(s, L);
    // phi(t) = t^L C(1/t)
    Poly phi((size_t)(kDeg / 64 + 1), 0);
    if (L != kDeg) return Poly();  // cannot happen for a nonzero MT19937 stream
    for (int k = 0; k <= L; ++k)
        if (bit(C, L - k)) phi[(size_t)(k >> 6)] |= 1ull << (k & 63);
    return phi;
}

struct Tables {
    Poly phi;                        // bits 0 .. kDeg
    std::vector<Poly> by_byte;       // (b(t) * t^kDeg) mod phi for the 256 bytes b
    std::vector<Poly> pow2;          // t^(2^k) mod phi, k = 0, 1, ...
    std::vector<std::vector<int32_t>> pow2_bits;  // their set bits, ascending

    // a mod phi for a of degree < 2 * kDeg (a has 2 * kPolyWords + 1 words)
    Poly reduce(Poly a) const {
        const int top = 2 * kDeg - 1;
        for (int r = (top - kDeg) / 8; r >= 0; --r) {
            const int64_t p = kDeg + 8 * (int64_t)r;  // bits [p, p + 8)
            uint32_t b = 0;
            for (int u = 0; u < 8; ++u)
                if (p + u < 64 * (int64_t)a.size() && bit(a, p + u)) b |= 1u << u;
            if (!b) continue;
            for (int u = 0; u < 8; ++u)
                if ((b >> u) & 1u) a[(size_t)((p + u) >> 6)] ^= 1ull << ((p + u) & 63);
            xor_shifted(a, by_byte[b], 8 * (int64_t)r);
        }
        a.resize(kPolyWords);
        return a;
    }
    Poly square(const Poly& g) const {
        Poly a((size_t)(2 * kPolyWords + 1), 0);
        for (int64_t i = 0; i < kDeg; ++i)
            if (bit(g, i)) a[(size_t)((2 * i) >> 6)] |= 1ull << ((2 * i) & 63);
        return reduce(a);
    }
};

inline Tables build_tables() {
    Tables t;
    // capacity for every k up to kMaxPow2: extending the lists never moves
    // the ones other threads read
    t.pow2.reserve(kMaxPow2 + 1);
    t.pow2_bits.reserve(kMaxPow2 + 1);
    t.phi = characteristic_polynomial();
    // low part of phi: t^kDeg = phi_low (mod phi)
    Poly low((size_t)kPolyWords, 0);
    for (int i = 0; i < kDeg; ++i)
        if (bit(t.phi, i)) low[(size_t)(i >> 6)] |= 1ull << (i & 63);
    // t^(kDeg + u) mod phi, u = 0..7
    Poly cur = low;
    std::vector<Poly> unit(8);
    for (int u = 0; u < 8; ++u) {
        unit[u] = cur;
        // cur *= t, reduced
        const int carry = bit(cur, kDeg - 1);
        for (int w = kPolyWords - 1; w >= 0; --w)
            cur[(size_t)w] = (cur[(size_t)w] << 1) | (w ? cur[(size_t)w - 1] >> 63 : 0);
        cur[(size_t)((kDeg) >> 6)] &= (1ull << (kDeg & 63)) - 1;  // drop bit kDeg
        if (carry)
            for (int w = 0; w < kPolyWords; ++w) cur[(size_t)w] ^= low[(size_t)w];
    }
    t.by_byte.assign(256, Poly((size_t)kPolyWords, 0));
    for (int b = 1; b < 256; ++b)
        for (int u = 0; u < 8; ++u)
            if ((b >> u) & 1)
                for (int w = 0; w < kPolyWords; ++w) t.by_byte[b][(size_t)w] ^= unit[u][(size_t)w];
    return t;
}

// The process-wide tables; pow2 is extended on demand up to k.
inline const Tables& tables(int k) {
    if (k > kMaxPow2) k = kMaxPow2;
    static std::mutex mu;
    static Tables* t = nullptr;
    std::lock_guard<std::mutex> lock(mu);
    if (!t) t = new Tables(build_tables());
    while ((int)t->pow2.size() <= k) {
        const int j = (int)t->pow2.size();
        Poly g;
        if ((int64_t(1) << (j < 62 ? j : 62)) < kDeg && j < 62) {  // t^(2^j) itself
            g.assign((size_t)kPolyWords, 0);
            const int64_t e = int64_t(1) << j;
            g[(size_t)(e >> 6)] |= 1ull << (e & 63);
        } else {
            g = t->square(t->pow2.back());
        }
        std::vector<int32_t> bits;
        for (int i = 0; i < kDeg; ++i)
            if (bit(g, i)) bits.push_back(i);
        t->pow2.push_back(std::move(g));
        t->pow2_bits.push_back(std::move(bits));
    }
    return *t;
}

// Host jump, the checker of the device one: the window 2^k words after w.
inline bool jump_pow2(const uint32_t* w, int k, uint32_t* out) {
    if (k < 0 || k > kMaxPow2) return false;
    const Tables& t = tables(k);
    if (t.phi.empty()) return false;
    std::vector<uint32_t> x((size_t)kStreamWords);
    stream_from(w, x.data(), kStreamWords);
    for (int j = 0; j < kN; ++j) {
        uint32_t acc = 0;
        for (int32_t i : t.pow2_bits[(size_t)k]) acc ^= x[(size_t)(i + j)];
        out[j] = acc;
    }
    return true;
}

}  // namespace mtj
