// queue_layout.cpp -- test-only check of the camera-bounce queue layout of
// quasi_amd/csrc/common.hpp (queue_seg_cap, queue_slots, queue_slot_of_path,
// queue_path_of_slot, queue_segment_paths) for pass sizes given on the
// command line, with the single-scene chunk, a batch's contiguous chunk and
// chunks of 64: every path gets one slot inside its segment's length, the
// segment lengths sum to the pass, the inverse returns the path, and the
// allocation (queue_slots_max) covers both layouts.  Prints "ok N" for N
// (size, chunk) cases checked or the first failure.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "common.hpp"

using namespace quasi_hip;

int main(int argc, char** argv) {
    int checked = 0;
    for (int a = 1; a < argc; ++a) {
        const uint32_t paths = uint32_t(std::strtoul(argv[a], nullptr, 10));
        for (const uint32_t chunk : {kQueueChunkScene, queue_contiguous_chunk(paths), 64u}) {
            if (chunk == 0u || chunk % 64u != 0u) { std::printf("fail %u: chunk %u\n", paths, chunk); return 1; }
            const uint32_t cap = queue_seg_cap(paths, chunk);
            const uint64_t slots = queue_slots(paths, chunk);
            if (slots < paths || slots > uint64_t(paths) + uint64_t(kQueueSegments) * chunk) {
                std::printf("fail %u/%u: %llu slots\n", paths, chunk, (unsigned long long)slots);
                return 1;
            }
            if (chunk != 64u && slots > queue_slots_max(paths)) {
                std::printf("fail %u/%u: %llu slots beyond the allocation\n", paths, chunk, (unsigned long long)slots);
                return 1;
            }
            uint64_t total = 0;
            std::vector<uint32_t> len(kQueueSegments);
            for (uint32_t s = 0; s < kQueueSegments; ++s) {
                len[s] = queue_segment_paths(paths, s, chunk);
                if (len[s] > cap) {
                    std::printf("fail %u/%u: segment %u holds %u > %u\n", paths, chunk, s, len[s], cap);
                    return 1;
                }
                total += len[s];
            }
            if (total != paths) {
                std::printf("fail %u/%u: segments hold %llu\n", paths, chunk, (unsigned long long)total);
                return 1;
            }
            std::vector<unsigned char> seen(slots, 0);
            for (uint32_t p = 0; p < paths; ++p) {
                const uint32_t i = queue_slot_of_path(p, cap, chunk);
                const uint32_t seg = i / cap;
                if (i >= slots || seg >= kQueueSegments || i - seg * cap >= len[seg] || seen[i]) {
                    std::printf("fail %u/%u: path %u slot %u\n", paths, chunk, p, i);
                    return 1;
                }
                seen[i] = 1;
                if (queue_path_of_slot(i, cap, chunk) != p) {
                    std::printf("fail %u/%u: slot %u back to %u, not %u\n", paths, chunk, i,
                                queue_path_of_slot(i, cap, chunk), p);
                    return 1;
                }
                // consecutive paths of a chunk sit in consecutive slots
                if (p % chunk != 0 && i != queue_slot_of_path(p - 1, cap, chunk) + 1) {
                    std::printf("fail %u/%u: path %u not next to %u\n", paths, chunk, p, p - 1);
                    return 1;
                }
            }
            ++checked;
        }
    }
    std::printf("ok %d\n", checked);
    return 0;
}
