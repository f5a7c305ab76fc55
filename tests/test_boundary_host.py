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


DATA = ("t # 0\n5 5 3 1\nv 0 1\nv 1 2\nv 2 2\nv 3 3\nv 4 1\n"
        "e 0 1 0\ne 1 0 0\ne 1 2 0\ne 2 2 0\ne 3 4 0\nt # -1\n")
QUERY = "t # 0\n2 1 1 1\nv 0 1\nv 1 2\ne 0 1 0\nt # -1\n"


def _read(driver, tmp_path, name, data):
    d = tmp_path / name
    d.write_bytes(data.encode())
    q = tmp_path / "q.g"
    q.write_text(QUERY)
    r = subprocess.run([driver, str(q), str(d), str(tmp_path / "out.txt")], capture_output=True, text=True)
    return r.returncode, r.stdout.splitlines()[:5]


def test_reader_token_forms(driver, tmp_path):
    """The reader takes tokens by scanf's rules (the reference reads them
    with fscanf, io/IO.cpp:53-97): spaces or tabs between tokens, none after
    't', a sign on a number, CRLF line ends, blank lines, trailing text after
    the numbers; an edge without a label gets label 0.  The graph read is
    the same."""
    rc, ref = _read(driver, tmp_path, "plain.g", DATA)
    assert rc == 0 and ref[0].startswith("n 5 adj 6 labels 3")
    variant = ("\n  t#0\r\n5 5 3 1\r\n\tv\t0   1\r\nv 1 2 extra\nv 2 +2\n\n   \nv 3 3\nv 4 1\n"
               "e 0 1\ne 1 0 7 junk\n  e 1 2 0\ne\t2 2 0\ne 3 4 -1\nt # -1\n")
    assert _read(driver, tmp_path, "variant.g", variant) == (0, ref)


@pytest.mark.parametrize("line", ["v 0", "v x 1", "v 1 1", "e 0", "e 0 x 1", "e 0 9 0", "w 0 1", "t # x"])
def test_reader_rejects_malformed_lines(driver, tmp_path, line):
    """A line that does not parse, a vertex out of order or an edge to a
    missing vertex ends the read with an error (the driver exits 4)."""
    if line.startswith("v"):
        bad = DATA.replace("v 0 1\n", line + "\n", 1)
    elif line.startswith("t"):  # the graph's own header (a later "t" line starts the next graph)
        bad = DATA.replace("t # 0\n", line + "\n", 1)
    else:
        bad = DATA.replace("e 3 4 0\n", "e 3 4 0\n" + line + "\n", 1)
    assert _read(driver, tmp_path, "bad.g", bad)[0] == 4
