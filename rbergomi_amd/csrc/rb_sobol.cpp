// rb_sobol.cpp -- Sobol' direction numbers (host), for the quasi-Monte Carlo Gaussian source.
//
// Replaces the point generator behind normalQMC_sample (rBergomi/qmc.cpp:26-41, i8_sobol of
// rBergomi/sobol.cpp), whose initial direction numbers are a table inside that source; this
// build derives its own: dimension 0 is the van der Corput sequence (every m_k = 1), dimension
// d >= 1 uses the d-th primitive polynomial over GF(2) in the order (degree, then the integer value
// of the coefficient word), with odd initial numbers m_1..m_s < 2^k drawn from a splitmix64
// stream keyed by (d, k), and the Bratley-Fox recurrence (ACM TOMS 659)
//   m_k = 2 a_1 m_{k-1} ^ 4 a_2 m_{k-2} ^ ... ^ 2^(s-1) a_(s-1) m_{k-s+1} ^ 2^s m_{k-s} ^ m_{k-s},
// p(x) = x^s + a_1 x^(s-1) + ... + a_(s-1) x + 1.  Direction integers V_k = m_k 2^(32-k),
// k = 1..32, so point i (0 <= i < 2^32) of dimension d is the XOR of V_k over the set bits
// k-1 of i, divided by 2^32.
#include <cstddef>
#include <cstdint>
#include <vector>

namespace rb {

namespace {

uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// (a * b) mod p over GF(2), deg p = d <= 30
uint64_t gf2_mulmod(uint64_t a, uint64_t b, uint64_t p, int d) {
    uint64_t r = 0;
    while (b) {
        if (b & 1) r ^= a;
        b >>= 1;
        a <<= 1;
        if (a >> d & 1) a ^= p;
    }
    return r;
}

uint64_t gf2_powx(uint64_t e, uint64_t p, int d) {  // x^e mod p
    uint64_t result = 1, base = d == 1 ? 1 : 2;       // x = 1 mod (x + 1)
    while (e) {
        if (e & 1) result = gf2_mulmod(result, base, p, d);
        base = gf2_mulmod(base, base, p, d);
        e >>= 1;
    }
    return result;
}

// x has multiplicative order 2^d - 1 modulo p (which also makes p irreducible)
bool primitive(uint64_t p, int d) {
    const uint64_t order = (uint64_t(1) << d) - 1;
    if (gf2_powx(order, p, d) != 1) return false;
    uint64_t n = order;
    for (uint64_t q = 2; q * q <= n; ++q) {
        if (n % q) continue;
        if (gf2_powx(order / q, p, d) == 1) return false;
        while (n % q == 0) n /= q;
    }
    return !(n > 1 && gf2_powx(order / n, p, d) == 1);
}

}  // namespace

// Primitive polynomials (coefficient words, bit s = x^s) in generation order, at least `count`.
std::vector<uint64_t> primitive_polynomials(int count) {
    std::vector<uint64_t> out;
    for (int d = 1; int(out.size()) < count && d <= 30; ++d)
        for (uint64_t mid = 0; mid < (uint64_t(1) << (d - 1)) && int(out.size()) < count; ++mid) {
            const uint64_t p = (uint64_t(1) << d) | (mid << 1) | 1;
            if (primitive(p, d)) out.push_back(p);
        }
    return out;
}

// V[d*32 + k-1] = direction integer k of dimension d, d < dims
void sobol_directions(int dims, std::vector<uint32_t>& V) {
    V.assign(std::size_t(dims) * 32, 0);
    for (int k = 1; k <= 32; ++k) V[k - 1] = uint32_t(1) << (32 - k);  // van der Corput
    if (dims < 2) return;
    const std::vector<uint64_t> polys = primitive_polynomials(dims - 1);
    for (int d = 1; d < dims; ++d) {
        const uint64_t p = polys[d - 1];
        int s = 0;
        while ((p >> (s + 1)) != 0) ++s;
        std::vector<uint64_t> m(33, 0);
        for (int k = 1; k <= s && k <= 32; ++k) {
            const uint64_t r = mix64(0x5B0B01ULL ^ (uint64_t(d) << 8) ^ uint64_t(k));
            m[k] = ((r % (uint64_t(1) << (k - 1))) << 1) | 1;  // odd, < 2^k
        }
        for (int k = s + 1; k <= 32; ++k) {
            uint64_t v = m[k - s] ^ (m[k - s] << s);
            for (int j = 1; j < s; ++j)
                if (p >> (s - j) & 1) v ^= m[k - j] << j;
            m[k] = v;
        }
        for (int k = 1; k <= 32; ++k) V[std::size_t(d) * 32 + k - 1] = uint32_t(m[k] << (32 - k));
    }
}

}  // namespace rb
