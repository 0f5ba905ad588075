// The reference-side use of the test_runner_t drop-in (include/pgap_test_runner.h),
// as run_gpu_thread makes it (gap_test_gpu.cu:236-248): numbers as mpz_t,
// one batch, results[i] in {0, 1}.  Test driver: `run_test_main ROUNDS`
// reads one hexadecimal number per line from stdin and prints each result
// on its own line; exit status 1 with a message when the runner fails.
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "pgap_test_runner.h"

int main(int argc, char **argv) {
    if (argc != 2) {
        fprintf(stderr, "usage: %s ROUNDS < numbers.hex\n", argv[0]);
        return 2;
    }
    const size_t rounds = strtoul(argv[1], nullptr, 10);
    std::vector<std::string> lines;
    for (std::string s; std::getline(std::cin, s);)
        if (!s.empty()) lines.push_back(s);
    std::vector<mpz_t *> tests;
    for (const std::string &s : lines) {
        mpz_t *z = (mpz_t *)malloc(sizeof(mpz_t));
        mpz_init_set_str(*z, s.c_str(), 16);
        tests.push_back(z);
    }
    std::vector<int> results(tests.size());
    int rc = 0;
    try {
        test_runner_t<> runner(tests.size(), rounds);
        runner.run_test(tests, results);
        for (int r : results) printf("%d\n", r);
    } catch (const std::exception &ex) {
        fprintf(stderr, "%s\n", ex.what());
        rc = 1;
    }
    for (mpz_t *z : tests) {
        mpz_clear(*z);
        free(z);
    }
    return rc;
}
