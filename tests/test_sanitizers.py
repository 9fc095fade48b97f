"""The CPU-side C of the build under AddressSanitizer and
UndefinedBehaviorSanitizer (gcc, host code only): the oracle over a few
steps under both CFL rules, the parameter-file readers on valid and
malformed files, and the state-file reader on valid, truncated and
corrupted files (its device copies stubbed with host memcpy). A run passes
when no sanitizer reports anything; malformed input must end in the
readers' own error exit (parameters) or error return (state files)."""
import os
import shutil
import struct
import subprocess

import pytest

from flow_amd import statefile as sf

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
SOURCES = [os.path.join(HERE, "hostcheck", "sanitize_host.c"),
           os.path.join(ROOT, "flow_amd", "csrc", "arch", "params.c"),
           os.path.join(ROOT, "flow_amd", "csrc", "hip", "state.c"),
           os.path.join(ROOT, "oracle", "hydro_omp3.c"),
           os.path.join(ROOT, "oracle", "arch_restated.c"),
           os.path.join(ROOT, "oracle", "driver_loop.c")]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="2")


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    if not shutil.which("gcc") or not os.path.exists(os.path.join(ROCM, "include", "hip")):
        pytest.skip("gcc or the HIP headers not found")
    out = str(tmp_path_factory.mktemp("san") / "sanitize_host")
    subprocess.run(["gcc", "-std=gnu99", "-g", "-O1", "-fno-omit-frame-pointer",
                    "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
                    "-fopenmp", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__",
                    f"-I{ROCM}/include", *SOURCES, "-lm", "-o", out],
                   check=True, capture_output=True, timeout=600)
    return out


def run(exe, *args, cwd=None):
    r = subprocess.run([exe, *map(str, args)], capture_output=True, text=True,
                       timeout=120, env=ENV, cwd=cwd)
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-3000:]
    return r


def test_oracle_is_clean(exe):
    r = run(exe, "oracle")
    assert r.returncode == 0, r.stderr[-2000:]
    assert r.stdout.count("oracle rule=") == 4


GOOD = """# comment
nx 64
 ny\t40
dt 1.0e-4
problem_0 density=0.125 energy=2.0 xpos=0.0 ypos=0.0 width=1.0 height=1.0
problem_1 density=1.0 energy=2.5 xpos=0.4 ypos=0.4 width=0.2 height=0.2
"""


@pytest.mark.parametrize("name,text,keys,ok", [
    ("good", GOOD, ("nx", "dt", "problem_1"), True),
    ("absent_spec", GOOD, ("ny", "dt", "problem_9"), True),
    ("missing_key", GOOD, ("iterations", "dt", "problem_0"), False),
    ("not_int", GOOD.replace("nx 64", "nx 6x4"), ("nx", "dt", "problem_0"), False),
    ("no_value", GOOD.replace("nx 64", "nx"), ("nx", "dt", "problem_0"), False),
    ("bad_pair", GOOD.replace("width=0.2", "width:0.2"), ("nx", "dt", "problem_1"), False),
    ("bad_number", GOOD.replace("width=0.2", "width=abc"), ("nx", "dt", "problem_1"), False),
    ("long_line", "nx " + "9" * 9000 + "\ndt 1\nproblem_0 a=1\n", ("nx", "dt", "problem_0"), False),
    ("long_key", "nx 3\ndt 1\nproblem_0 " + "k" * 3000 + "=1\n", ("nx", "dt", "problem_0"), True),
    ("many_keys", "nx 3\ndt 1\nproblem_0 " + " ".join(f"k{i}=1" for i in range(300)) + "\n",
     ("nx", "dt", "problem_0"), False),
    ("empty", "", ("nx", "dt", "problem_0"), False),
    ("binary", "\x00\x01\xff" * 500, ("nx", "dt", "problem_0"), False),
])
def test_params_readers(exe, tmp_path, name, text, keys, ok):
    p = tmp_path / f"{name}.params"
    p.write_bytes(text.encode("latin-1"))
    r = run(exe, "params", p, *keys)
    if ok:
        assert r.returncode == 0, r.stderr[-2000:]
    else:
        assert r.returncode == 1, (r.returncode, r.stderr[-2000:])


def test_state_round_trip_and_corruptions(exe, tmp_path):
    p = tmp_path / "ok.state"
    w = run(exe, "state-write", p, 20, 13)
    assert w.returncode == 0 and " rc=0 " in w.stdout
    r = run(exe, "state-read", p, 20, 13)
    assert " rc=0 step=7 " in r.stdout
    assert w.stdout.split("checksum=")[1] == r.stdout.split("checksum=")[1]
    assert sf.read(p)["ny"] == 13
    good = p.read_bytes()
    cases = {
        "empty": b"",
        "magic": b"X" + good[1:],
        "version": good[:8] + struct.pack("<I", 9) + good[12:],
        "huge_nx": good[:16] + struct.pack("<i", 1 << 30) + good[20:],
        "negative_ny": good[:20] + struct.pack("<i", -5) + good[24:],
        "pad": good[:24] + struct.pack("<i", 3) + good[28:],
        "layout": good[:72 + 16] + b"Z" + good[72 + 17:],
        "count": good[:72 + 24] + struct.pack("<Q", 1 << 62) + good[72 + 32:],
        "name": good[:72] + b"rhq" + good[75:],
    }
    for cut in (1, 40, 71, 72 + 31, 72 + 32 + 8, len(good) - 1):
        cases[f"cut{cut}"] = good[:cut]
    for name, data in cases.items():
        q = tmp_path / f"{name}.state"
        q.write_bytes(data)
        r = run(exe, "state-read", q, 20, 13)
        assert r.returncode == 0, (name, r.stderr[-2000:])
        assert " rc=-1 " in r.stdout, (name, r.stdout)
    # a valid file of another mesh is refused, not read past its end
    r = run(exe, "state-read", p, 21, 13)
    assert " rc=-1 " in r.stdout
