// gap_sieve_hip -- writes combined_sieve unknown files with the device sieve
// of libpgap (pgap_sieve): for every m in [mstart, mstart + minc) coprime to
// d, the offsets x in [-SL, SL] for which m*P#/d + x has no prime factor <=
// max_prime, in the plain (default), run-length (--rle) or bit-compressed
// (--bitcompressed) line format.
//
// Reference counterpart: combined_sieve --save (combined_sieve.cpp:74-157,
// writers :1261-1445), argument names of Args::argparse (gap_common.cpp:
// 805-1099).  Only the unknown file is produced; the reference's range /
// range_stats database rows are out of scope.
#include <getopt.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../../include/pgap.h"
#include "hostbn.h"
#include "unknowns.h"

using namespace pgap;

static void usage(const char *name) {
    printf("Usage: %s -p <p> -d <d> --mstart <m> --minc <n> --sieve-length <SL> --max-prime <MP in millions>\n", name);
    printf("   or: %s -u <P>_<D>_<mstart>_<minc>_s<SL>_l<MP>M.txt\n", name);
    printf("  --rle | --bitcompressed   output format (default plain)\n");
    printf("  --method1                 name the file .m1.txt (same content)\n");
    printf("  -q                        quiet\n");
}

int main(int argc, char **argv) {
    static struct option opts[] = {{"p", required_argument, 0, 'p'},
                                   {"d", required_argument, 0, 'd'},
                                   {"mstart", required_argument, 0, 1},
                                   {"minc", required_argument, 0, 2},
                                   {"sieve-length", required_argument, 0, 4},
                                   {"max-prime", required_argument, 0, 5},
                                   {"unknown-filename", required_argument, 0, 'u'},
                                   {"save", no_argument, 0, 7},
                                   {"save-unknowns", no_argument, 0, 7},
                                   {"rle", no_argument, 0, 13},
                                   {"bitcompressed", no_argument, 0, 15},
                                   {"method1", no_argument, 0, 8},
                                   {"quiet", no_argument, 0, 'q'},
                                   {"help", no_argument, 0, 'h'},
                                   {0, 0, 0, 0}};
    UnknownConfig u;
    int format = 0, verbose = 2, c, idx = 0;
    std::string err;
    bool from_name = false;
    while ((c = getopt_long(argc, argv, "qhp:d:u:", opts, &idx)) >= 0) {
        switch (c) {
            case 'h': usage(argv[0]); return 0;
            case 'q': --verbose; break;
            case 'p': u.p = (uint32_t)atoi(optarg); break;
            case 'd': u.d = (uint32_t)atoi(optarg); break;
            case 1: u.mstart = strtoull(optarg, nullptr, 10); break;
            case 2: u.minc = strtoull(optarg, nullptr, 10); break;
            case 4: u.sieve_length = (uint32_t)atoi(optarg); break;
            case 5: u.max_prime = strtoull(optarg, nullptr, 10) * 1000000ull; break;
            case 7: break;
            case 8: u.method1 = true; break;
            case 13: format = 1; break;
            case 15: format = 2; break;
            case 'u':
                if (!parse_unknown_filename(optarg, u, err)) {
                    fprintf(stderr, "%s\n", err.c_str());
                    return 1;
                }
                from_name = true;
                break;
            default: usage(argv[0]); return 1;
        }
    }
    if (!u.p || !u.d || !u.mstart || !u.minc || !u.sieve_length || !u.max_prime) {
        usage(argv[0]);
        return 1;
    }
    if (!from_name) u.path = gen_unknown_fn(u);
    Words K;
    if (!primorial_over(u.p, u.d, K)) {
        fprintf(stderr, "d=%u must divide %u# with P#/d > 1\n", u.d, u.p);
        return 1;
    }
    pgap_engine *e = nullptr;
    pgap_options opt{};
    int fam = -1;
    if (pgap_create(&opt, &e) != PGAP_OK || pgap_family(e, u.p, u.d, u.sieve_length, &fam) != PGAP_OK) {
        fprintf(stderr, "engine: %s\n", pgap_last_error(e));
        pgap_destroy(e);
        return 1;
    }
    const size_t slash = u.path.find_last_of('/');
    if (slash != std::string::npos) mkdir(u.path.substr(0, slash).c_str(), 0755);
    FILE *out = fopen(u.path.c_str(), "wb");
    if (!out) {
        fprintf(stderr, "can't write %s\n", u.path.c_str());
        pgap_destroy(e);
        return 1;
    }
    if (verbose >= 1) printf("Saving unknowns to '%s'\n", u.path.c_str());
    BitLayout bl;
    if (format == 2) bl.build(K, u.p, u.d, u.sieve_length);

    std::vector<uint64_t> ms;
    for (uint64_t mi = 0; mi < u.minc; ++mi)
        if (gcd64(u.mstart + mi, u.d) == 1) ms.push_back(u.mstart + mi);
    const size_t batch = 4096, SL = u.sieve_length;
    std::vector<int32_t> pv(batch * SL), nx(batch * SL);
    std::vector<uint32_t> np(batch), nn(batch);
    size_t total = 0;
    int status = 0;
    for (size_t off = 0; off < ms.size() && !status; off += batch) {
        const size_t nb = std::min(batch, ms.size() - off);
        if (pgap_sieve(e, fam, ms.data() + off, nb, u.max_prime, pv.data(), np.data(), nx.data(), nn.data()) !=
            PGAP_OK) {
            fprintf(stderr, "sieve: %s\n", pgap_last_error(e));
            status = 1;
            break;
        }
        for (size_t j = 0; j < nb; ++j) {
            std::vector<int32_t> a(pv.begin() + j * SL, pv.begin() + j * SL + np[j]);
            std::vector<int32_t> b(nx.begin() + j * SL, nx.begin() + j * SL + nn[j]);
            total += a.size() + b.size();
            const std::string line = format == 0   ? plain_line(ms[off + j], a, b)
                                     : format == 1 ? rle_line(ms[off + j], a, b)
                                                   : bitcompressed_line(ms[off + j], a, b, bl);
            fwrite(line.data(), 1, line.size(), out);
        }
    }
    fclose(out);
    if (verbose >= 2 && !status)
        printf("\t%zu m's, %zu unknowns (%.2f per m)\n", ms.size(), total, ms.empty() ? 0.0 : (double)total / ms.size());
    pgap_destroy(e);
    return status;
}
