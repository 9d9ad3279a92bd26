"""The host code that reads untrusted input (packet and header parsing in
csrc/prr_parse.cpp, the movie reader csrc/prr_mov.cpp, the scene encoder
csrc/synth.c that builds the seed packets) built with AddressSanitizer and
UndefinedBehaviorSanitizer (g++/gcc, no HIP) and fuzzed on the CPU by
tests/fuzz_host.cpp: mutated scene packets, synth random frames and movies
written by this package. Any sanitizer report fails the run."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "proresraw_rs_amd", "csrc")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
       "-g", "-O1"]


@pytest.fixture(scope="module")
def fuzz_binary(tmp_path_factory):
    out = tmp_path_factory.mktemp("san")
    objs = []
    for src, cc, std in (("synth.c", "gcc", "-std=gnu11"), ("prr_parse.cpp", "g++", "-std=c++17"),
                         ("prr_mov.cpp", "g++", "-std=c++17")):
        obj = str(out / (src + ".o"))
        subprocess.run([cc, *SAN, std, "-c", os.path.join(CSRC, src), "-o", obj], check=True)
        objs.append(obj)
    exe = str(out / "fuzz_host")
    # The sanitizer runtimes linked statically, so the binary needs nothing
    # loaded before them.
    subprocess.run(["g++", *SAN, "-static-libasan", "-static-libubsan", "-std=c++17",
                    os.path.join(ROOT, "tests", "fuzz_host.cpp"), *objs, "-o", exe, "-lpthread",
                    "-lm"], check=True)
    return exe


def seed_movies(tmp_path):
    import proresraw_rs_amd as P  # noqa: F401  (packets come from the encoder)
    from proresraw_rs_amd import mov as M
    from proresraw_rs_amd import synth as S
    paths = []
    pk = S.encode_scene(S.scene(64, 32, 0, seed=3, version=0, alignment=1, threads=1))
    for i, (n, fps) in enumerate(((3, (24000, 1001)), (1, (25, 1)), (5, (30, 1)))):
        p = tmp_path / f"seed{i}.mov"
        p.write_bytes(M.clip_movie([pk] * n, fps=fps))
        paths.append(str(p))
    return paths


def test_fuzzed_host_parsers_under_sanitizers(fuzz_binary, tmp_path):
    movies = seed_movies(tmp_path)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([fuzz_binary, "20000", "12345", *movies], capture_output=True, text=True,
                       env=env, timeout=600)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0, r.stderr[-4000:]
    assert "done: 20000 iterations" in r.stdout
