// scene_cornell.hpp -- the reference's default scene as a triangle mesh.
#pragma once

#include <cstdint>
#include <vector>

namespace quasi_hip {

// The Cornell box of src/quasi/scene/cornell_box.hpp:95-134: 5 walls, the
// ceiling light quad and two rotated 5-face boxes, 16 quads in the reference
// order.  Quad i (origin o, edges u, v) becomes triangles 2i = (o, o+u, o+u+v)
// and 2i+1 = (o, o+u+v, o+v), which keep the quad's normal direction cross(u, v).
// `aspect` is accepted for symmetry with make_cornell_box; the geometry does
// not depend on it.
void make_cornell_box(float aspect, std::vector<float>& verts, std::vector<uint32_t>& indices,
                      std::vector<float>& albedo, std::vector<float>& emission);

}  // namespace quasi_hip
