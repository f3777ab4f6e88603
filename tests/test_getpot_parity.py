"""The drop-in driver's input.dat parser (mesobasic_amd/host/InputParams.cpp)
against the reference's own GetPot (src/utils/GetPot, compiled in place
into oracle/_ref/getpot_check by oracle/Makefile.ref): on a corpus of inputs
-- section nesting, repeated keys, several assignments on a line, quoting,
comments, malformed lines, and numbers that GetPot converts through
sscanf("%lf") (ints through double) -- both print the same sections,
variables and typed lookups, byte for byte.  Skipped where the harness is not
built (it needs the reference tree)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GETPOT = os.path.join(ROOT, "oracle", "_ref", "getpot_check")
MESO = os.path.join(ROOT, "mesobasic_amd", "host", "meso_amd")

pytestmark = pytest.mark.skipif(not os.path.exists(GETPOT), reason="oracle/_ref/getpot_check not built")

CORPUS = {
    "grammar": """
# leading comment
[Domain]
  nx = 64   # trailing comment
  ny=32
  nz =16
[Time]
  nstep = 2.0e3
[LBApp]
  app = true
  G = -5.5
  [./inner]
    a = 1
  [../other]
    b = 2
[]
  root = 3
[Domain]
  nx = 65
""",
    "numbers": """
[N]
  a = 12abc
  b = 3.99
  c = -0
  d = 1e400
  e = nan
  f = inf
  g = 0x1A
  h = .5
  i = 7.
  j = +5
  k = 1.5e-310
  l = 2147483648
  m = -2147483649
  n = 1e-5000
  o = 3.7e1
  p = -2.5
  q = 9.99999999999999999
  r = 0.1e1
  s = -1e-320
  t = 1.7976931348623157e308
  u = 00012
  v = 1,5
""",
    "strings": """
[S]
  q1 = 'a b c'
  q2 = ''
  q3 = "dq"
  q4 = 'a # b'
  q5 = x'y'z
  two = 1 three = 3
  tab\t=\t9
""",
    "malformed": """
dangling =
= lone
[Domain
]
[[x]]
[ A ]
  k = 1
[a/b]
  c = 2
[./d]
  e = 3
[../f]
  g = 4
[]
  h = 5
""",
    "crlf": "[Domain]\r\n  nx = 8\r\n  ny = 9\r\n[Time]\r\n  nstep = 10",
    # an empty '#' line swallows the next line (the character after '#' is
    # always read as part of the comment)
    "comment_quirks": "[D]\n#\n  nx = 5\n  ny = 6 #c\n#x\n  nz = 7\n  a#b = 8\n",
    "escapes": "[E]\n  a = x\\ y\n  b = 'it\\'s'\n  c = back\\\\slash\n  d = 1\\\n  e = 2\n",
    "spacing": "[F]\n  a= b\n  c =d\n  e = \n  g\n  =\n  h = = i\n  j == 3\n",
    "section_eq": "[a=b]\n  k = 1\n[c]\n  m = 2\n",
    "eof_eq": "[G]\n k =",
    "empty": "",
    "comments_only": "# a\n# b = 3\n   # [x]\n",
}

MISSING = ["i:Missing/k:7", "d:Missing/k:2.5", "s:Missing/k:dflt", "i:Domain/none:-3"]


def _run(exe, path, queries):
    args = [exe, path] + queries if exe == GETPOT else [exe, "--query", path] + queries
    r = subprocess.run(args, capture_output=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return r.stdout


@pytest.mark.parametrize("name", sorted(CORPUS))
def test_parser_matches_getpot(tmp_path, name):
    if not os.path.exists(MESO):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "mesobasic_amd", "host")], check=True)
    path = tmp_path / "input.dat"
    path.write_bytes(CORPUS[name].encode())
    listing = _run(GETPOT, str(path), [])
    keys = [ln.split(" = ", 1)[0][4:] for ln in listing.decode().splitlines() if ln.startswith("var ")]
    queries = [f"{t}:{k}:{d}" for k in keys for t, d in (("i", "-7"), ("d", "0.25"), ("s", "none"))]
    queries += MISSING
    want = _run(GETPOT, str(path), queries)
    got = _run(MESO, str(path), queries)
    assert got == want, f"{name}:\n--- GetPot\n{want.decode()}\n--- driver\n{got.decode()}"


def test_expressions_are_refused(tmp_path):
    # GetPot would evaluate ${...}; the driver's reader refuses the file
    # instead of reading it differently
    if not os.path.exists(MESO):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "mesobasic_amd", "host")], check=True)
    path = tmp_path / "input.dat"
    path.write_text("[Domain]\n  nx = 64\n  ny = ${Domain/nx}\n")
    r = subprocess.run([MESO, "--query", str(path)], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "${...}" in r.stderr
