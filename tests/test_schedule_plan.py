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


@pytest.fixture(scope="module")
def split_cost(tmp_path_factory):
    exe = tmp_path_factory.mktemp("split") / "split_cost"
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "subway_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "split_cost.cpp"), "-o", str(exe)], check=True)

    def run(degrees, after, min_cost):
        np_ = [0]
        for d in degrees:
            np_.append(np_[-1] + d)
        line = f"{len(degrees)} {after} {min_cost}\n" + " ".join(map(str, np_)) + "\n"
        out = subprocess.run([str(exe)], input=line, capture_output=True, text=True, check=True).stdout
        return [int(x) for x in out.split()]
    return run


def test_low_degree_unit_pieces_halve_to_the_floor(split_cost):
    """1000 vertices of degree 1 (cost 4 each: the edge plus three for the
    vertex), nothing behind them: pieces take half of the cost left, 2000,
    1000, 500, 250, then the floor of 200 (50 vertices) ends the halving."""
    ends = split_cost([1] * 1000, 0, 200)
    sizes = [b - a for a, b in zip([0] + ends, ends)]
    assert sizes[:4] == [500, 250, 125, 62] and sum(sizes) == 1000 and ends[-1] == 1000
    assert all(s >= 50 for s in sizes[:-1]) and max(sizes[4:]) <= 62


def test_work_behind_a_unit_keeps_it_whole(split_cost):
    # half of the cost still to come exceeds the unit: one piece
    assert split_cost([1] * 100, 10_000, 1) == [100]


def test_pieces_follow_cost_not_vertex_count(split_cost):
    # a heavy vertex (degree 400, cost 403) then 100 of degree 1 (cost 4):
    # the first half of the cost (503 of 803 / 2 = 401, at least one vertex)
    # is the heavy vertex alone
    ends = split_cost([400] + [1] * 100, 0, 1)
    assert ends[0] == 1 and ends[-1] == 101
    assert all(b > a for a, b in zip([0] + ends, ends))
