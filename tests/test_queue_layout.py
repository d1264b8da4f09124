"""The camera-bounce queue layout (quasi_amd/csrc/common.hpp) on the CPU.

The wavefront and single-launch engines deal a pass's path ids to the 64 queue
segments in chunks of consecutive ids, round robin: 1024-path chunks for a
single scene (every segment holds a sample of the whole image), one contiguous
chunk per segment for a batch.  tests/native/queue_layout.cpp checks, for each
pass size and both chunk choices (and chunks of 64), that the segment lengths
add up to the pass and fit the segment capacity, that every path has one slot
inside its segment's length (no two paths share one), that the inverse mapping
returns the path, that a chunk's paths sit in consecutive slots (a wave of the
camera bounce takes one pixel's samples) and that the allocation covers both
layouts.  Sizes: tiny passes, chunk and segment-round boundaries,
a 1080p frame at 1 and 64 samples, and the largest automatic pass (2^27).
"""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    out = tmp_path_factory.mktemp("queue") / "queue_layout"
    subprocess.run(["g++", "-O2", "-std=c++20", "-I", str(ROOT / "quasi_amd" / "csrc"),
                    str(ROOT / "tests" / "native" / "queue_layout.cpp"), "-o", str(out)], check=True)
    return out


def test_queue_layout_is_a_permutation(probe):
    sizes = [1, 2, 63, 64, 65, 127, 128, 4031, 4095, 4096, 4097, 8191, 8192, 8193, 64 * 64 * 3 + 5, 100_000,
             262_144, 262_145, 1920 * 1080, 1920 * 1080 * 64 // 8 + 17]
    res = subprocess.run([str(probe), *map(str, sizes)], capture_output=True, text=True, check=True).stdout
    assert res.strip() == f"ok {3 * len(sizes)}", res


def test_queue_layout_largest_pass(probe):
    # 2^27 paths: one whole 64-spp 1080p frame fits below it
    res = subprocess.run([str(probe), str(1 << 27), str(1920 * 1080 * 64)], capture_output=True, text=True,
                         check=True).stdout
    assert res.strip() == "ok 6", res
