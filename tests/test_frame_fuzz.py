"""The host .frame parser under AddressSanitizer and UBSan (CPU only): the
parser reads untrusted files, so truncated and corrupted frames must end in
a status code, never in an out-of-bounds access.  frame_io.cpp and the
harness are compiled with g++ alone (no device code).  Leak reports are off:
the parser's host thread pool keeps one zstd context per thread for the
life of the process."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
ZSTD = "/usr/lib/x86_64-linux-gnu/libzstd.so.1"


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.exists(ZSTD), reason="needs g++ and libzstd")
def test_frame_parser_fuzz_asan(pvvc, tmp_path):
    imgs = pvvc.synth_stream("c0", n_frames=2, seed=7, threads=8)
    frames = []
    for k, b in enumerate(imgs):
        p = tmp_path / f"f{k}.frame"
        p.write_bytes(b)
        frames.append(str(p))
    exe = tmp_path / "frame_fuzz"
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
           "-I/opt/conda/include", f"-I{ROOT / 'include'}", f"-I{ROOT / 'patchvvc-plus_amd' / 'csrc'}",
           str(ROOT / "patchvvc-plus_amd" / "csrc" / "frame_io.cpp"), str(ROOT / "tests" / "fuzz" / "frame_fuzz.cpp"),
           "-o", str(exe), ZSTD, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1", UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([str(exe), "6000", *frames], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    assert "ERROR" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    parsed, rejected = (int(x) for x in r.stdout.split()[1::2])
    assert parsed > 1000 and rejected > 1000
