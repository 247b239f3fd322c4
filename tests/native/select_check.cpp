// Host check of treecorr_amd/csrc/tc_select.h against libstdc++'s std::nth_element:
// the same comparisons and moves must leave the same permutation.  Built and run
// by tests/test_select.py.
#include <algorithm>
#include <cstdio>
#include <random>
#include <vector>
#include "../../treecorr_amd/csrc/tc_select.h"

int main() {
    std::mt19937_64 rng(12345);
    long bad = 0, total = 0;
    for (int trial = 0; trial < 4000; ++trial) {
        long n = 1 + long(rng() % (trial < 3000 ? 64 : 3000));
        int nvals = (trial % 3 == 0) ? 3 : (trial % 3 == 1 ? int(n) : 1000000);
        std::vector<double> key(n);
        for (long i = 0; i < n; ++i) key[i] = double(rng() % nvals) * 0.25;
        if (trial % 7 == 0) std::sort(key.begin(), key.end());
        if (trial % 11 == 0) std::sort(key.rbegin(), key.rend());
        long first = 0, last = n;
        long nth = (first + last) / 2;
        std::vector<int> a(n), b(n);
        for (long i = 0; i < n; ++i) a[i] = b[i] = int(i);
        tc_select::KeyLess cmp{key.data()};
        std::nth_element(a.begin(), a.begin() + nth, a.end(), [&](int x, int y) { return key[x] < key[y]; });
        tc_select::nth_element(b.data(), first, nth, last, cmp);
        ++total;
        if (a != b) ++bad;
    }
    std::printf("select_check: %ld of %ld permutations differ\n", bad, total);
    return bad == 0 ? 0 : 1;
}
