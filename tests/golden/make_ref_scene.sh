#!/bin/sh
# Regenerates tests/golden/ref_cornell_scene.json: the reference's own
# make_cornell_box() and camera::get_ray() (src/quasi/scene/cornell_box.hpp,
# camera.hpp), compiled from the reference sources by oracle/Makefile.
set -e
here=$(dirname "$0")
make -s -C "$here/../../oracle" ref
cp "$here/../../oracle/_ref/ref_scene.json" "$here/ref_cornell_scene.json"
