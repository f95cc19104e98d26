"""Mutation check of the response oracle's pins (TEST INFRASTRUCTURE, CPU only).

Each mutation below changes one token of oracle/oracle_response.cpp (a link
index, a sign inside a sinc argument, a spherical-harmonic coefficient, a sign
in a TDI combination); the mutant oracle is built in a temporary directory and
the oracle pin tests (tests/test_oracle_response_pins.py, tests/test_oracle_pins.py)
run against it. Every mutant must make at least one pin fail.

    python oracle/mutation_check.py
"""
import re
import shutil
import subprocess
import sys
import tempfile
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
SRC = ROOT / "oracle"
MUT = [
    ("link index kn3 -> kn1 in G12", "g.G12 = common * n3Hn3 * sinc(pre * (1. - kn3)) * e12;", "g.G12 = common * n3Hn3 * sinc(pre * (1. - kn1)) * e12;"),
    ("sinc sign in G23", "g.G23 = common * n1Hn1 * sinc(pre * (1. - kn1)) * e23;", "g.G23 = common * n1Hn1 * sinc(pre * (1. + kn1)) * e23;"),
    ("sYlm (2,2) coefficient 64 -> 32", "if (l == 2 && m == 2) fac = std::sqrt(5.0 / (64.0 * PI))", "if (l == 2 && m == 2) fac = std::sqrt(5.0 / (32.0 * PI))"),
    ("sYlm (2,1) (1 + c) -> (1 - c)", "else if (l == 2 && m == 1) fac = std::sqrt(5.0 / (16.0 * PI)) * st * (1.0 + c);", "else if (l == 2 && m == 1) fac = std::sqrt(5.0 / (16.0 * PI)) * st * (1.0 - c);"),
    ("T channel sign G32 - G23 -> G23 - G32", "cplx T = INVSQRT6 * (g.G21 - g.G12 + g.G32 - g.G23 + g.G13 - g.G31);", "cplx T = INVSQRT6 * (g.G21 - g.G12 + g.G23 - g.G32 + g.G13 - g.G31);"),
]
tmp = Path(tempfile.mkdtemp(prefix="oracle_mutants_"))
survivors = []
for name, a, b in MUT:
    d = tmp / re.sub(r"\W+", "_", name)
    d.mkdir(parents=True)
    for f in SRC.glob("*.cpp"):
        shutil.copy(f, d)
    shutil.copy(SRC / "oracle.h", d)
    s = (d / "oracle_response.cpp").read_text()
    assert s.count(a) == 1, name
    (d / "oracle_response.cpp").write_text(s.replace(a, b))
    lib = d / "libmut.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-w", "-fopenmp", "-shared", "-o", str(lib),
                    *[str(d / x) for x in ("oracle_phenomhm.cpp", "oracle_response.cpp", "oracle_build.cpp", "oracle_spline.cpp")]], check=True)
    r = subprocess.run([sys.executable, "-c", f"""
import sys, ctypes; sys.path.insert(0, {str(ROOT)!r})
import oracle; oracle._lib = ctypes.CDLL({str(lib)!r})
import pytest; sys.exit(pytest.main(['-q', '-p', 'no:cacheprovider', {str(ROOT / 'tests' / 'test_oracle_response_pins.py')!r}, {str(ROOT / 'tests' / 'test_oracle_pins.py')!r}]))
"""], capture_output=True, text=True, cwd=str(ROOT))
    tail = [l for l in r.stdout.splitlines() if "passed" in l or "failed" in l][-1:]
    failed = [l.split("::")[-1].split(" ")[0] for l in r.stdout.splitlines() if l.startswith("FAILED")]
    print(f"{name}: exit {r.returncode}; {tail}; failing: {failed}", flush=True)
    if r.returncode == 0:
        survivors.append(name)
shutil.rmtree(tmp, ignore_errors=True)
print("every mutant caught" if not survivors else f"mutants not caught: {survivors}")
sys.exit(1 if survivors else 0)
