"""Host check of the learned schedule's choice of in-place visits
(choose_resident_visits in subway_amd/csrc/schedule.hpp, which the engine
calls every global iteration): slices of whole id ranges with enough active
edges are visited where they lie, and when the two buffers hold overlapping
slices (moved there in different iterations) the overlap is visited once."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = tmp_path_factory.mktemp("plan") / "resident_choice"
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "subway_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "resident_choice.cpp"), "-o", str(exe)], check=True)

    def run(bounds, active, slices, nb=2, min_bytes=1, eb=4):
        P = len(bounds) - 1
        (lo0, hi0), (lo1, hi1) = slices
        line = " ".join(map(str, [P, *bounds, *active, lo0, hi0, lo1, hi1, nb, min_bytes, eb]))
        out = subprocess.run([str(exe)], input=line + "\n", capture_output=True, text=True, check=True).stdout.split()
        v0, v1, ov, *done = map(int, out)
        return (bool(v0), bool(v1)), ov, done
    return run


BOUNDS = [0, 10, 20, 30, 40, 50, 60, 70, 80]   # eight id ranges


def test_disjoint_slices_both_visited(checker):
    visit, ov, done = checker(BOUNDS, [5] * 8, [(0, 40), (40, 80)])
    assert visit == (True, True) and ov == 0 and done == [1] * 8


def test_overlapping_slices_visited_once(checker):
    """Buffer 0 holds ranges 0-3, buffer 1 ranges 2-5 (an overlap of two
    ranges): buffer 0's slice is visited, buffer 1's is skipped and counted,
    and ranges 4-5 stay for the transfers of the iteration."""
    visit, ov, done = checker(BOUNDS, [5] * 8, [(0, 40), (20, 60)])
    assert visit == (True, False) and ov == 1
    assert done == [1, 1, 1, 1, 0, 0, 0, 0]


def test_overlap_left_to_the_qualifying_slice(checker):
    """Buffer 0's slice has too few active edges: buffer 1's overlapping
    slice is then visited, not skipped."""
    visit, ov, done = checker(BOUNDS, [0, 0, 0, 0, 9, 9, 0, 0], [(0, 40), (20, 60)], min_bytes=64)
    assert visit == (False, True) and ov == 0
    assert done == [0, 0, 1, 1, 1, 1, 0, 0]


def test_minimum_volume_and_alignment(checker):
    # 7 active edges * 4 bytes < 32: not worth a visit
    visit, _, _ = checker(BOUNDS, [1] * 8, [(0, 70), (0, 0)], min_bytes=32)
    assert visit == (False, False)
    # a slice that does not start on a range bound is never visited
    visit, _, done = checker(BOUNDS, [5] * 8, [(5, 40), (40, 80)])
    assert visit == (False, True) and done == [0, 0, 0, 0, 1, 1, 1, 1]


def test_one_buffer(checker):
    visit, ov, _ = checker(BOUNDS, [5] * 8, [(0, 40), (0, 80)], nb=1)
    assert visit == (True, False) and ov == 0


@pytest.fixture(scope="module")
def tail_cap(tmp_path_factory):
    exe = tmp_path_factory.mktemp("tail") / "tail_cap"
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "subway_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "tail_cap.cpp"), "-o", str(exe)], check=True)

    def run(rows):
        line = "\n".join(" ".join(map(str, r)) for r in rows)
        out = subprocess.run([str(exe)], input=line + "\n", capture_output=True, text=True, check=True).stdout
        return [int(x) for x in out.split()]
    return run


def test_whole_units_shrink_towards_the_end(tail_cap):
    """Units of whole ranges, built greedily with the cap whole_unit_cap gives
    at each unit's start: full buffers while plenty remains, then halves, and
    the last unit is one range (8 ranges of 10 per buffer of 80, 60 ranges)."""
    B, R, n = 80, 10, 60
    caps, units, left = [], [], n * R
    i = 0
    while i < n:
        cap = tail_cap([(B, R, left, 1)])[0]
        size = 0
        while i < n and size + R <= cap:
            size += R
            i += 1
        units.append(size)
        left -= size
    assert units == [80, 80, 80, 80, 80, 80, 60, 30, 10, 10, 10]
    assert sum(units) == n * R


def test_tail_floor_keeps_small_units_whole(tail_cap):
    # below the floor nothing is split; the opening range always fits
    assert tail_cap([(80, 10, 100, 50)]) == [50]
    assert tail_cap([(80, 10, 30, 50)]) == [50]
    assert tail_cap([(80, 90, 30, 1)]) == [90]
    assert tail_cap([(80, 10, 1000, 1)]) == [80]
