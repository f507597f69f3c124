"""Config (include/fdm/config.h) and unixbench_score (include/fdm/unixbench_score.h)
against the reference's src/config.cpp and src/unixbench_score.cpp.

The query harness tests/cpp/config_probe.cpp is built against this build's
headers and, where the reference checkout is present, against the reference's
own sources (oracle/Makefile -> oracle/_ref/config_probe).  Both read the same
INI text, command-line overrides and queries; their answers must be identical.
The parsing rules the reference's loader implies (src/config.cpp:78-120: the
section word is sscanf("[%s]") less its last character, '#'/';' comment lines,
';' cuts a line, key and value are the first two words split on " =\\t",
lines before any section are ignored) are also pinned by explicit expected
answers, so the test keeps its value where the reference build is absent.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OWN = os.path.join(ROOT, "tests", "cpp", "config_probe")
REF = os.path.join(ROOT, "oracle", "_ref", "config_probe")

INI = """top = 1
[ns]
nx = 15 ; interior points per axis
Re=100
# comment = 3
; also = 4
dt\t0.01
u0 = 2 extra words
  nz = 31
ny =
[ plot ]
vtk 1
interval = 1.5e2
[ns ]
key = in_n
[x]]
k = v
"""

QUERIES = [
    ("i ns nx 32", "15"), ("i ns nz 32", "31"), ("d ns Re 1", "100"), ("d ns dt 0.001", "0.01"),
    ("s ns u0 1", "2"), ("i ns ny 7", "7"),                      # "ny =" has no value: default
    ("i ns comment 0", "0"), ("i ns also 0", "0"), ("i ns top 5", "5"),   # comments, pre-section line
    ("i plot vtk 0", "0"), ("i plo vtk 0", "1"),                 # "[ plot ]" names section "plo"
    ("i plo interval 100", "1"), ("d plo interval 100", "150"),  # int from "1.5e2" stops at '.'
    ("s n key none", "in_n"),                                   # "[ns ]" names section "n"
    ("s x] k none", "v"),                                       # "[x]]" names section "x]"
    ("d ns x1 -3.141592653589793", "-3.1415926535897931"),       # default round trip
    ("S ns nx", "15"), ("S ns missing", "throw"), ("S nosuch nx", "throw"),
    ("D ns dt", "0.01"), ("I plo vtk", "1"),
    ("i ns steps 1", "40"), ("s solver datatype double", "float"),   # from the command line
    ("u 1 2 4 8 16 32", "2.8284271247461898"),                 # geometric mean of the fastest 4 of 6
    ("u 9 3 6", "4.2426406871192848"),                         # sorts first: sqrt(3 * 6)
    ("u 7 1", "1"), ("u 5", None), ("u", "0"),                  # 2 keep 1; 1 keeps none (0/0); empty -> 0
]
ARGS = ["--ns:steps=40", "--solver:datatype=float", "not-an-override", "--noColon=1", "--ns:bad"]


def build_own():
    if not os.path.exists(OWN):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp"), OWN])
    return OWN


def run(binary, ini_path, queries):
    r = subprocess.run([binary, str(ini_path)] + ARGS, input="\n".join(queries) + "\np\n", capture_output=True,
                       text=True, timeout=30)
    assert r.returncode == 0, r.stderr
    return r.stdout.splitlines()


def test_config_rules_pinned(tmp_path):
    ini = tmp_path / "t.ini"
    ini.write_text(INI)
    out = run(build_own(), ini, [q for q, _ in QUERIES])
    for (q, want), got in zip(QUERIES, out):
        if want is None:
            assert got.lstrip("-") == "nan", (q, got)
        else:
            assert got == want, (q, got, want)
    dump = out[len(QUERIES):]
    assert dump[0] == "[n]" and dump[-1] == "."
    assert "[ns]" in dump and "nx = 15" in dump and "steps = 40" in dump and "datatype = float" in dump


def test_config_missing_file_gives_defaults(tmp_path):
    out = run(build_own(), tmp_path / "absent.ini", ["i ns nx 32", "d ns Re 1", "S ns nx"])
    assert out[:3] == ["32", "1", "throw"]


@pytest.mark.skipif(not os.path.exists(REF), reason="reference Config build (oracle/_ref/config_probe) absent")
def test_config_matches_reference_build(tmp_path):
    ini = tmp_path / "t.ini"
    ini.write_text(INI + "[ns]\nRe = 250\n\n\r\n[]\nempty = section\n")
    queries = [q for q, _ in QUERIES] + ["d ns Re 1", "i ns nx 0"]   # "p" dumps every section, "[]" included
    assert run(build_own(), ini, queries) == run(REF, ini, queries)
