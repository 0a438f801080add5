// CPU check of the cone passes' link-range plan (tsr_amd/csrc/link_ranges.h) with record
// totals far above 2^32: every accepted range stays below the record cap (so its 32-bit
// offsets cannot wrap), the accepted ranges cover the links once and in order, merging is
// on whenever there is more than one range, and a single link over the cap is refused.
// Per-link record counts are piecewise constant over blocks of 1000 links, so that the
// record total of any range is exact in O(1).
#include <cstdint>
#include <cstdio>
#include <vector>

#include "../../tsr_amd/csrc/link_ranges.h"

namespace {

struct Counts {
    std::vector<uint64_t> perLink, prefix;   // per block of 1000 links
    uint64_t upTo(uint64_t i) const {        // records of links [0, i)
        const uint64_t b = i / 1000;
        return prefix[b] + (i % 1000) * (b < perLink.size() ? perLink[b] : 0);
    }
    uint64_t range(uint32_t a, uint32_t n) const { return upTo((uint64_t)a + n) - upTo(a); }
};

Counts pattern(uint32_t links, uint64_t base, uint64_t spike, uint32_t period) {
    Counts c;
    const uint32_t blocks = (links + 999) / 1000 + 1;
    c.perLink.resize(blocks);
    c.prefix.resize(blocks + 1);
    for (uint32_t b = 0; b < blocks; ++b) c.perLink[b] = base + (b % period == period - 1 ? spike : 0);
    for (uint32_t b = 0; b < blocks; ++b) c.prefix[b + 1] = c.prefix[b] + 1000 * c.perLink[b];
    return c;
}

int fails = 0;
void expect(bool ok, const char *what) {
    if (!ok) {
        std::printf("FAIL %s\n", what);
        ++fails;
    }
}

// runs the plan as the cone pass does; returns the number of accepted ranges
int run(uint32_t links, uint32_t chunkLinks, uint64_t maxRecords, double estimate, const Counts &c, bool &refused) {
    tsr::LinkRanges r(links, chunkLinks, maxRecords, estimate);
    uint32_t covered = 0;
    int accepted = 0, guard = 0;
    uint64_t sum = 0;
    refused = false;
    while (!r.done()) {
        if (++guard > 10000000) {
            expect(false, "plan terminates");
            break;
        }
        const uint32_t l0 = r.begin;
        const uint32_t n = r.next();
        expect(n >= 1 && (uint64_t)l0 + n <= links, "range inside the links");
        if (n < links) expect(r.merging, "merging before a partial range");
        const uint64_t total = c.range(l0, n);
        bool impossible = false;
        if (!r.accept(n, total, impossible)) {
            if (impossible) {
                refused = true;
                return accepted;
            }
            expect(r.merging, "merging after a split");
            continue;
        }
        expect(total < maxRecords, "accepted range below the record cap");
        expect(total < (1ull << 32), "accepted range fits 32-bit offsets");
        expect(l0 == covered, "ranges in order, without gaps");
        covered += n;
        sum += total;
        ++accepted;
        r.finished(n, total);
    }
    expect(covered == links, "every link covered");
    expect(sum == c.upTo(links), "record total preserved");
    if (accepted > 1) expect(r.merging, "merging with several ranges");
    return accepted;
}

}  // namespace

int main() {
    bool refused = false;
    const uint64_t cap = 1536ull << 20;   // the cone passes' record cap
    // 40M links averaging ~150 records (6e9 in total, above 2^32), spikes of 20000 records
    // per link every 50 blocks; no estimate: the first range is split until it fits
    Counts a = pattern(40000000u, 140, 20000, 50);
    expect(a.upTo(40000000u) > (1ull << 32), "total above 2^32");
    int k = run(40000000u, 1u << 24, cap, 0.0, a, refused);
    expect(!refused && k > 3, "split plan");
    std::printf("no estimate: %d ranges for %.3g records\n", k, (double)a.upTo(40000000u));
    // with a (low) estimate of the records per link
    k = run(40000000u, 1u << 24, cap, 100.0, a, refused);
    expect(!refused && k > 3, "estimated plan");
    // fewer links than one chunk but more records than the cap: still merged (several ranges)
    Counts b = pattern(10000000u, 400, 0, 1);
    k = run(10000000u, 1u << 24, cap, 0.0, b, refused);
    expect(!refused && k >= 3, "one chunk of links, several ranges");
    // one range only: no merging
    Counts d = pattern(1000000u, 10, 0, 1);
    tsr::LinkRanges one(1000000u, 1u << 24, cap, 0.0);
    one.next();
    expect(!one.merging, "a single range is not merged");
    k = run(1000000u, 1u << 24, cap, 0.0, d, refused);
    expect(k == 1 && !refused, "single range");
    // a single link above the cap is refused, after splitting down to it
    Counts e = pattern(4000u, 1, 2000000000ull, 3);   // block 2 (links 2000..2999): 2e9 records each
    run(4000u, 1u << 24, cap, 0.0, e, refused);
    expect(refused, "a link above the cap is refused");
    std::printf(fails ? "FAILED %d\n" : "OK\n", fails);
    return fails ? 1 : 0;
}
