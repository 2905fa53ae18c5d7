// Checks the saturating-offset logic the device scans and the candidate fill share (curvedspacesim_amd/csrc/
// scan_math.h): a scan whose total passes INT32_MAX keeps every offset non-negative and non-decreasing, no write
// lands outside [0, cap), and exactly the sources that lose candidates are flagged. Prints "ok" or the failure.
#include <cstdio>
#include <vector>

#include "scan_math.h"

using namespace csg;

static int check(const std::vector<long long>& counts, long long cap) {
    const int n = (int)counts.size();
    std::vector<int> off(n + 1);
    long long run = 0;
    for (int i = 0; i < n; ++i) {  // the scan kernels: 64-bit running sums, saturated stores
        off[i] = saturateOffset(run);
        run += counts[i];
    }
    off[n] = saturateOffset(run);
    std::vector<char> seen(cap, 0);
    long long trueBase = 0;
    for (int i = 0; i < n; ++i) {
        if (off[i] < 0 || off[i] > off[i + 1]) return printf("offset %d not monotonic: %d %d\n", i, off[i], off[i + 1]), 1;
        long long written = 0;
        for (long long p = 0; p < counts[i]; ++p)
            if (candidateSlotValid(off[i], p, cap)) {
                long long slot = off[i] + p;
                if (slot < 0 || slot >= cap) return printf("write outside the buffer: %lld\n", slot), 1;
                if (seen[slot]) return printf("slot %lld written twice\n", slot), 1;
                seen[slot] = 1;
                ++written;
            }
        const bool lost = written < counts[i];
        if (lost != candidatesOverflow(off[i], counts[i], cap))
            return printf("source %d: lost %d but flagged %d\n", i, (int)lost, (int)candidatesOverflow(off[i], counts[i], cap)), 1;
        if (counts[i] > 0 && !lost && trueBase != off[i]) return printf("source %d: offset %d, true %lld\n", i, off[i], trueBase), 1;
        trueBase += counts[i];
    }
    return 0;
}

int main() {
    const long long big = kOffsetMax / 2 + 7;
    int bad = 0;
    bad |= check({3, 5, 0, 2}, 10);                  // fits exactly
    bad |= check({3, 5, 0, 2, 4}, 10);               // the last source overflows
    bad |= check({big, big, 5, 0, 3}, 64);           // totals pass INT32_MAX: offsets saturate instead of wrapping
    bad |= check({1, big, big, big, 2, 0}, 1000);    // several times past INT32_MAX
    bad |= check({0, 0, 0}, 1);                      // nothing to place
    if (saturateOffset(-1) != -1 || saturateOffset((long long)kOffsetMax + 1) != kOffsetMax) bad = 1;
    printf(bad ? "failed\n" : "ok\n");
    return bad;
}
