"""The C++ boundary classes on the CPU: Graph's CSR fields are plain arrays
as in the reference (graph/Graph.h:99-119), so reference-style code that
passes them as unsigned* compiles, and IO's four output overloads write the
reference's text (io/IO.cpp:129-142, 182-232)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "gpsm_amd", "csrc", "host")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("boundary") / "boundary_driver")
    subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-I", HOST,
                    os.path.join(ROOT, "tests", "boundary_driver.cpp"), os.path.join(HOST, "Graph.cpp"),
                    os.path.join(HOST, "IO.cpp"), "-o", exe], check=True)
    return exe


def test_graph_arrays_and_io_output_overloads(driver, tmp_path):
    # 5 vertices; 0-1 twice (both directions) and a self-loop 2-2: the CSR
    # keeps one undirected edge per pair and drops the loop
    data = tmp_path / "data.g"
    data.write_text("t # 0\n5 5 3 1\nv 0 1\nv 1 2\nv 2 2\nv 3 3\nv 4 1\n"
                    "e 0 1 0\ne 1 0 0\ne 1 2 0\ne 2 2 0\ne 3 4 0\nt # -1\n")
    query = tmp_path / "query.g"
    query.write_text("t # 0\n2 1 1 1\nv 0 1\nv 1 2\ne 0 1 0\nt # -1\n")
    out = tmp_path / "ans.txt"
    r = subprocess.run([driver, str(query), str(data), str(out)], capture_output=True, text=True, check=True)
    lines = r.stdout.splitlines()
    # rows: 0:[1] 1:[0,2] 2:[1] 3:[4] 4:[3]; checksum = labels + sum col*(v+1)
    row_sum = 1 * 1 + (0 + 2) * 2 + 1 * 3 + 4 * 4 + 3 * 5
    assert lines[0] == "n 5 adj 6 labels 3 checksum %d" % (9 + row_sum)
    assert lines[1:4] == ["label 1 x2", "label 2 x2", "label 3 x1"]
    assert lines[4] == "deg0 1 off0 0 deg_last 1"
    assert lines[5] == "copy checksum %d" % (9 + row_sum)
    text = out.read_text()
    expect = ("query graph:0    data graph:0\n" + "=" * 60 + "\n"
              "(0, 4) (1, 0) (2, 2) \n"
              # IO::output(final_result, ...): column id_map[j] holds query vertex j
              "(0, 3) (1, 1) (2, 2) \n(0, 1) (1, 3) (2, 2) \n\n\n\n"
              "\n\n\n")
    assert text == expect
