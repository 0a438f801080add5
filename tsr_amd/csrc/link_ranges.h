// Link ranges of one cone pass (host logic only, no HIP): the links are traversed in
// ascending ranges so that one range's records fit one sort and 32-bit record offsets.
//
// A range is proposed (next), counted on the device, and then either accepted or split:
// the counting traversal sums the range's records in 64 bits, so a range whose total
// reaches maxRecords is halved and traversed again before any 32-bit offset is formed
// (a range of one link that still reaches it cannot be processed).  After an accepted
// range, the next range is sized to about 3/4 of maxRecords at the records per link seen.
// Results of more than one range are merged per leaf in range order; `merging` turns on
// before the first range that does not cover every link, and stays on.
#pragma once

#include <algorithm>
#include <cstdint>

namespace tsr {

struct LinkRanges {
    uint32_t links = 0, chunkLinks = 0;
    uint64_t maxRecords = 0;
    uint32_t rangeLinks = 0;   // proposed length of the next range
    uint32_t begin = 0;        // first link of the next range
    bool merging = false;

    // perLinkEstimate: records per link estimated beforehand (0 = none)
    LinkRanges(uint32_t links_, uint32_t chunkLinks_, uint64_t maxRecords_, double perLinkEstimate)
        : links(links_), chunkLinks(chunkLinks_), maxRecords(maxRecords_), rangeLinks(chunkLinks_) {
        if (perLinkEstimate > 0.0) size(perLinkEstimate);
    }
    bool done() const { return begin >= links; }
    // length of the next range [begin, begin + n)
    uint32_t next() {
        const uint32_t n = std::min<uint32_t>(rangeLinks, links - begin);
        if (n < links) merging = true;
        return n;
    }
    // after counting [begin, begin + n): false if the range must be split and counted again
    // (`impossible` is set when a single link reaches maxRecords)
    bool accept(uint32_t n, uint64_t total, bool &impossible) {
        impossible = false;
        if (total < maxRecords) return true;
        if (n < 2) {
            impossible = true;
            return false;
        }
        merging = true;
        rangeLinks = n / 2;
        return false;
    }
    // after an accepted range was processed
    void finished(uint32_t n, uint64_t total) {
        size(std::max(1.0, (double)total / std::max<uint32_t>(n, 1)));
        begin += n;
    }

private:
    void size(double perLink) {
        rangeLinks = (uint32_t)std::max(1.0, std::min((double)chunkLinks, 0.75 * (double)maxRecords / perLink));
    }
};

}  // namespace tsr
