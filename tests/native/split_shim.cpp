// Host entry points over quasi_amd/csrc/bvh_split.hpp for tests/test_bvh_split.py:
// the split of a triangle into BVH references and the split threshold,
// compiled for the host from the builders' own header.
#include "bvh_split.hpp"

using namespace quasi_hip;

extern "C" {

// Triangle i = tri[9i .. 9i+9) (three corners); out[24i ..] receives up to
// four boxes (lo.xyz, hi.xyz each), count[i] how many.
void ss_split(const float* tri, uint32_t n, float T, const int32_t* may_split, float* out, uint32_t* count) {
    for (uint32_t i = 0; i < n; ++i) {
        PieceBox pc[kMaxPieces];
        const float* v = tri + 9 * size_t(i);
        const uint32_t c = split_triangle(v, v + 3, v + 6, T, may_split[i] != 0, pc);
        count[i] = c;
        for (uint32_t j = 0; j < c; ++j)
            for (int k = 0; k < 3; ++k) {
                out[24 * size_t(i) + 6 * j + k] = pc[j].lo[k];
                out[24 * size_t(i) + 6 * j + 3 + k] = pc[j].hi[k];
            }
    }
}

float ss_threshold(const float* lo, const float* hi, uint32_t n) { return split_threshold(lo, hi, n); }

}  // extern "C"
