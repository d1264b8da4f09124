// Host entry points over quasi_amd/csrc/bvh_quant.hpp for tests/test_bvh_quant.py:
// the quantiser and the quantised slab test that the traversal kernels run,
// compiled for the host from the same header.
#include "bvh_quant.hpp"

using namespace quasi_hip;

extern "C" {

// nodes: n 128-byte BVH4 nodes (32 floats each); out: n quantised nodes (16
// floats each); ok: one flag per node.
void qs_quantize(const float* nodes, uint32_t n, float* out, int32_t* ok) {
    for (uint32_t i = 0; i < n; ++i)
        ok[i] = quantize_node4(reinterpret_cast<const float4*>(nodes) + 8 * size_t(i),
                               reinterpret_cast<float4*>(out) + 4 * size_t(i))
                    ? 1
                    : 0;
}

// For node i of qnodes and ray i (origin o[3i..], inverse direction
// inv[3i..], interval [tmin, tmax[i]]): tn/tf of the four children.
void qs_slab(const float* qnodes, const float* o, const float* inv, float tmin, const float* tmax, uint32_t n,
             float* tn, float* tf) {
    for (uint32_t i = 0; i < n; ++i) {
        const float4* q = reinterpret_cast<const float4*>(qnodes) + 4 * size_t(i);
        float a[4], b[4];
        slab_quant4(q[0], q[1], q[2], o[3 * i], o[3 * i + 1], o[3 * i + 2], inv[3 * i], inv[3 * i + 1],
                    inv[3 * i + 2], tmin, tmax[i], a, b);
        for (int k = 0; k < 4; ++k) {
            tn[4 * i + k] = a[k];
            tf[4 * i + k] = b[k];
        }
    }
}

}  // extern "C"
