// unknowns.h -- the combined_sieve unknown-file contract, host side (C++17):
// the file-name grammar, format detection and line parsing for the plain,
// run-length and bit-compressed formats, plus writers for all three.
//
// Reference counterparts: Args::argparse -u handling (gap_common.cpp:877-925),
// gen_unknown_fn (:732-749), guess_compression (:752-802), BitArrayHelper
// (:1124-1181), parse_unknown_line (:1186-1359) and the writers of
// save_unknowns_method2 (combined_sieve.cpp:1261-1445).
//
// Line formats (distances are positive; prev and next in outward order):
//   plain (0)          m : -A +B | -p1 -p2 ... | +n1 +n2 ...
//   run-length (1)     m : -A +B | <pairs> | <pairs>     pair = 48+(d>>7), 48+(d&127),
//                                                       d = delta from the previous distance
//   bit-compressed (2) m : U BYTES || <raw>             one bit per offset x in [-SL, SL]
//                      with gcd(K, x) = 1 and gcd(m*K + x, d) = 1, low bit first, 7 bits
//                      per byte with the high bit set; bit 1 = composite.
#pragma once
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "hostbn.h"

namespace pgap {

struct UnknownConfig {
    uint32_t p = 0, d = 0;
    uint64_t mstart = 0, minc = 0;
    uint32_t sieve_length = 0;
    uint64_t max_prime = 0;
    bool method1 = false;
    std::string path;  // file actually opened
};

inline uint64_t gcd64(uint64_t a, uint64_t b) {
    while (b) {
        uint64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

// "<P>_<D>_<mstart>_<minc>_s<SL>_l<MP>M[.m1].txt", optionally with a directory;
// a bare name is looked up under unknowns/ as the reference does.
inline bool parse_unknown_filename(const std::string &arg, UnknownConfig &c, std::string &err) {
    const size_t slash = arg.find_last_of('/');
    const std::string base = slash == std::string::npos ? arg : arg.substr(slash + 1);
    const std::string dir = slash == std::string::npos ? "unknowns" : arg.substr(0, slash);
    unsigned long long P, D, ms, mi, SL, MP;
    char tail[32] = {0};
    int used = 0;
    if (std::count(base.begin(), base.end(), '_') != 5 ||
        sscanf(base.c_str(), "%llu_%llu_%llu_%llu_s%llu_l%lluM%n", &P, &D, &ms, &mi, &SL, &MP, &used) != 6) {
        err = "unknown-file name must be <P>_<D>_<mstart>_<minc>_s<SL>_l<maxprime>M.txt: " + base;
        return false;
    }
    snprintf(tail, sizeof tail, "%s", base.c_str() + used);
    if (strcmp(tail, ".txt") != 0 && strcmp(tail, ".m1.txt") != 0) {
        err = "unknown-file name must end in M.txt or M.m1.txt: " + base;
        return false;
    }
    c.p = (uint32_t)P;
    c.d = (uint32_t)D;
    c.mstart = ms;
    c.minc = mi;
    c.sieve_length = (uint32_t)SL;
    c.max_prime = MP * 1000000ull;
    c.method1 = strcmp(tail, ".m1.txt") == 0;
    c.path = dir + "/" + base;
    return true;
}

inline std::string gen_unknown_fn(const UnknownConfig &c) {
    return "unknowns/" + std::to_string(c.p) + "_" + std::to_string(c.d) + "_" + std::to_string(c.mstart) + "_" +
           std::to_string(c.minc) + "_s" + std::to_string(c.sieve_length) + "_l" +
           std::to_string(c.max_prime / 1000000) + "M" + (c.method1 ? ".m1" : "") + ".txt";
}

// Everything a bit-compressed line needs about K and d, computed once.
struct BitLayout {
    uint32_t SL = 0, D = 0;
    std::vector<int32_t> coprime_x;      // x in [-SL, SL] with gcd(K, x) = 1, increasing
    std::vector<uint32_t> d_primes;      // prime factors of d
    std::vector<uint32_t> k_mod_dprime;  // K mod q for each q | d

    void build(const Words &K, uint32_t P, uint32_t D_, uint32_t SL_) {
        SL = SL_;
        D = D_;
        std::vector<uint32_t> kprimes;  // primes dividing K: p <= P, p does not divide d
        d_primes.clear();
        for (uint32_t p : primes_upto(P)) (D % p == 0 ? d_primes : kprimes).push_back(p);
        k_mod_dprime.clear();
        for (uint32_t q : d_primes) k_mod_dprime.push_back(mod_small(K, q));
        std::vector<uint8_t> bad(2 * (size_t)SL + 1, 0);
        for (uint32_t p : kprimes) {
            // offsets divisible by p: index SL + x with x = 0 mod p
            for (int64_t x = -(int64_t)(SL - SL % p); x <= (int64_t)SL; x += p) bad[(size_t)(x + SL)] = 1;
        }
        coprime_x.clear();
        for (int64_t x = -(int64_t)SL; x <= (int64_t)SL; ++x)
            if (!bad[(size_t)(x + SL)] && x != 0) coprime_x.push_back((int32_t)x);
    }

    // gcd(m*K + x, d) == 1
    bool coprime_to_d(uint64_t m, int32_t x) const {
        for (size_t i = 0; i < d_primes.size(); ++i) {
            const uint64_t q = d_primes[i];
            const uint64_t mk = (m % q) * k_mod_dprime[i] % q;
            const uint64_t xr = (uint64_t)(((int64_t)x % (int64_t)q + (int64_t)q) % (int64_t)q);
            if ((mk + xr) % q == 0) return false;
        }
        return true;
    }
};

// Format of a file from its first line: 2 if it has " || ", else 0 when
// both sides are signed decimal lists, else 1.
inline int detect_format(const std::string &line) {
    if (line.find(" || ") != std::string::npos) return 2;
    const size_t a = line.find(" | ");
    if (a == std::string::npos) return -1;
    const size_t b = line.find(" | ", a + 3);
    if (b == std::string::npos) return -1;
    auto plain_side = [](const std::string &s, char sign) {
        for (size_t i = 0; i < s.size(); ++i) {
            const char ch = s[i];
            if (ch == ' ' || ch == '\n' || ch == '\r') continue;
            if (ch == sign && (i == 0 || s[i - 1] == ' ')) continue;
            if (ch < '0' || ch > '9') return false;
        }
        return true;
    };
    const std::string l = line.substr(a + 3, b - a - 3), r = line.substr(b + 3);
    if (!l.empty() && l[0] != '-') return 1;
    return plain_side(l, '-') && plain_side(r, '+') ? 0 : 1;
}

// Parse one line.  Returns false (with err) on a malformed line instead of
// the reference's assert (gap_common.cpp:1201-1358).
inline bool parse_unknown_line(const std::string &line, int format, const BitLayout *bl, uint64_t &m,
                               std::vector<int32_t> &prev, std::vector<int32_t> &next, std::string &err) {
    prev.clear();
    next.clear();
    const size_t colon = line.find(" : ");
    if (colon == std::string::npos) {
        err = "missing ' : '";
        return false;
    }
    char *end = nullptr;
    m = strtoull(line.c_str(), &end, 10);
    if (end != line.c_str() + colon) {
        err = "bad m";
        return false;
    }
    const char *p = line.c_str() + colon + 3;
    if (format == 2) {
        if (!bl) {
            err = "bit-compressed line needs K";
            return false;
        }
        unsigned long long total = 0, nbytes = 0;
        int used = 0;
        if (sscanf(p, "%llu %llu ||%n", &total, &nbytes, &used) != 2 || used == 0 || p[used] != ' ') {
            err = "bad bit-compressed header";
            return false;
        }
        const unsigned char *raw = (const unsigned char *)p + used + 1;
        const size_t avail = line.size() - (size_t)((const char *)raw - line.c_str());
        if (avail < nbytes) {
            err = "truncated bit-compressed line";
            return false;
        }
        size_t bit = 0;
        for (int32_t x : bl->coprime_x) {
            if (!bl->coprime_to_d(m, x)) continue;
            if (bit / 7 >= nbytes) {
                err = "bit-compressed line shorter than its coprime set";
                return false;
            }
            const unsigned v = (raw[bit / 7] >> (bit % 7)) & 1u;
            ++bit;
            if (v == 0) (x < 0 ? prev : next).push_back(x < 0 ? -x : x);
        }
        if ((bit + 6) / 7 != nbytes || prev.size() + next.size() != total) {
            err = "bit-compressed counts disagree";
            return false;
        }
        std::reverse(prev.begin(), prev.end());
        return true;
    }
    long long a = 0, b = 0;
    int used = 0;
    if (sscanf(p, "-%lld +%lld |%n", &a, &b, &used) != 2 || used == 0 || a < 0 || b < 0) {
        err = "bad header";
        return false;
    }
    p += used;
    if (*p != ' ') {
        err = "expected ' ' after '|'";
        return false;
    }
    ++p;
    const char *lend = line.c_str() + line.size();
    auto read_side = [&](std::vector<int32_t> &out, long long cnt, char sign) -> bool {
        if (format == 1) {
            int32_t acc = 0;
            for (long long k = 0; k < cnt; ++k) {
                if (p + 2 > lend) return false;
                acc += 128 * ((unsigned char)p[0] - 48) + ((unsigned char)p[1] - 48);
                p += 2;
                out.push_back(acc);
            }
            return true;
        }
        for (long long k = 0; k < cnt; ++k) {
            while (p < lend && *p == ' ') ++p;
            if (p >= lend || *p != sign) return false;
            ++p;
            char *e2 = nullptr;
            const long v = strtol(p, &e2, 10);
            if (e2 == p || v <= 0) return false;
            out.push_back((int32_t)v);
            p = e2;
        }
        return true;
    };
    if (!read_side(prev, a, '-')) {
        err = "bad prev list";
        return false;
    }
    // separator " | "
    while (p < lend && *p == ' ') ++p;
    if (p >= lend || *p != '|') {
        err = "missing second '|'";
        return false;
    }
    ++p;
    if (p < lend && *p == ' ') ++p;
    if (!read_side(next, b, '+')) {
        err = "bad next list";
        return false;
    }
    return true;
}

// ---- writers (the combined_sieve output side) ----

inline std::string plain_line(uint64_t m, const std::vector<int32_t> &prev, const std::vector<int32_t> &next) {
    std::string s = std::to_string(m) + " : -" + std::to_string(prev.size()) + " +" + std::to_string(next.size()) +
                    " |";
    for (int32_t d : prev) s += " -" + std::to_string(d);
    s += " |";
    for (int32_t d : next) s += " +" + std::to_string(d);
    s += "\n";
    return s;
}

inline std::string rle_line(uint64_t m, const std::vector<int32_t> &prev, const std::vector<int32_t> &next) {
    std::string s = std::to_string(m) + " : -" + std::to_string(prev.size()) + " +" + std::to_string(next.size()) +
                    " | ";
    auto enc = [&](const std::vector<int32_t> &v) {
        int32_t last = 0;
        for (int32_t d : v) {
            const int32_t delta = d - last;
            last = d;
            s += (char)(48 + (delta >> 7));
            s += (char)(48 + (delta & 127));
        }
    };
    enc(prev);
    s += " | ";
    enc(next);
    s += "\n";
    return s;
}

inline std::string bitcompressed_line(uint64_t m, const std::vector<int32_t> &prev, const std::vector<int32_t> &next,
                                      const BitLayout &bl) {
    std::vector<uint8_t> is_unknown(2 * (size_t)bl.SL + 1, 0);
    for (int32_t d : prev) is_unknown[(size_t)(bl.SL - d)] = 1;
    for (int32_t d : next) is_unknown[(size_t)(bl.SL + d)] = 1;
    std::string raw;
    unsigned char byte = 0x80;
    int bit = 0;
    for (int32_t x : bl.coprime_x) {
        if (!bl.coprime_to_d(m, x)) continue;
        if (!is_unknown[(size_t)(x + (int32_t)bl.SL)]) byte |= (unsigned char)(1u << bit);
        if (++bit == 7) {
            raw += (char)byte;
            byte = 0x80;
            bit = 0;
        }
    }
    if (bit) raw += (char)byte;
    return std::to_string(m) + " : " + std::to_string(prev.size() + next.size()) + " " + std::to_string(raw.size()) +
           " || " + raw + "\n";
}

}  // namespace pgap
