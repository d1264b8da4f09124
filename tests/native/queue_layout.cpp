// queue_layout.cpp -- test-only check of the camera-bounce queue layout of
// quasi_amd/csrc/common.hpp (queue_seg_cap, queue_slots, queue_slot_of_path,
// queue_path_of_slot, queue_segment_paths) for pass sizes given on the
// command line: every path gets one slot inside its segment's length, the
// segment lengths sum to the pass, and the inverse returns the path.  Prints
// "ok N" for N sizes checked or the first failure.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "common.hpp"

using namespace quasi_hip;

int main(int argc, char** argv) {
    int checked = 0;
    for (int a = 1; a < argc; ++a) {
        const uint32_t paths = uint32_t(std::strtoul(argv[a], nullptr, 10));
        const uint32_t cap = queue_seg_cap(paths);
        const uint64_t slots = queue_slots(paths);
        if (slots < paths || slots > uint64_t(paths) + kQueueSegments * kQueueChunk) {
            std::printf("fail %u: %llu slots\n", paths, (unsigned long long)slots);
            return 1;
        }
        uint64_t total = 0;
        std::vector<uint32_t> len(kQueueSegments);
        for (uint32_t s = 0; s < kQueueSegments; ++s) {
            len[s] = queue_segment_paths(paths, s);
            if (len[s] > cap) { std::printf("fail %u: segment %u holds %u > %u\n", paths, s, len[s], cap); return 1; }
            total += len[s];
        }
        if (total != paths) { std::printf("fail %u: segments hold %llu\n", paths, (unsigned long long)total); return 1; }
        std::vector<unsigned char> seen(slots, 0);
        for (uint32_t p = 0; p < paths; ++p) {
            const uint32_t i = queue_slot_of_path(p, cap);
            const uint32_t seg = i / cap;
            if (i >= slots || seg >= kQueueSegments || i - seg * cap >= len[seg] || seen[i]) {
                std::printf("fail %u: path %u slot %u\n", paths, p, i);
                return 1;
            }
            seen[i] = 1;
            if (queue_path_of_slot(i, cap) != p) {
                std::printf("fail %u: slot %u back to %u, not %u\n", paths, i, queue_path_of_slot(i, cap), p);
                return 1;
            }
            // consecutive paths of a chunk sit in consecutive slots
            if (p % kQueueChunk != 0 && i != queue_slot_of_path(p - 1, cap) + 1) {
                std::printf("fail %u: path %u not next to %u\n", paths, p, p - 1);
                return 1;
            }
        }
        ++checked;
    }
    std::printf("ok %d\n", checked);
    return 0;
}
