"""The device headers' sampler, tree traversal, lane-group split and facet
lists, built for
the host with g++ (GPU qualifiers stripped from copies of the headers, the
tree built by the library's own host code) and checked against the oracle's
sampleTriangle and against brute force: tests/host/headers_check.cpp."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "tetwild_amd", "csrc")


def host_copy(src, dst):
    s = open(src).read()
    s = s.replace("#include <hip/hip_runtime.h>", "#include <cmath>\n#include <cstdint>\n#include <algorithm>")
    s = re.sub(r"__attribute__\(\(address_space\(3\)\)\)", "", s)
    s = re.sub(r"\b__host__\b|\b__device__\b", "", s)
    s = s.replace("__forceinline__", "inline")
    s = s.replace("__syncthreads();", "").replace("threadIdx.x", "0").replace("blockDim.x", "1")
    s = s.replace("__clz(", "__builtin_clz(")
    # device-only helpers the checks do not use
    s = re.sub(r"inline int64_t xcd_block\(\)\s*\{.*?\n\}", "", s, flags=re.S)
    if os.path.basename(src) == "twg_bvh.hpp":
        # one-lane groups: a ballot is the lane's own predicate
        s = s.replace("__builtin_amdgcn_fence(__ATOMIC_RELEASE, \"wavefront\");", "")
        s = s.replace("__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, \"wavefront\");", "")
        s = s.replace("__builtin_amdgcn_wave_barrier();", "").replace("__popcll(", "__builtin_popcountll(")
        s = ("struct uint4 { unsigned x, y, z, w; };\n"
             "inline unsigned long long __ballot(bool b) { return b ? 1ull : 0ull; }\n"
             "template <class T> inline T __shfl(T v, int, int = 64) { return v; }\n"
             "template <class T> inline T __shfl_xor(T v, int, int = 64) { return v; }\n") + s
    open(dst, "w").write(s)


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("hdr")
    for h in ("twg_math.hpp", "twg_geom.hpp", "twg_bvh.hpp"):
        host_copy(os.path.join(CSRC, h), str(d / h))
    env = open(os.path.join(CSRC, "envelope.hip")).read()
    i = env.index("inline uint64_t spread3")
    j = env.index("}  // namespace\n}  // namespace twg", i)
    (d / "build_part.hpp").write_text("namespace twg {\n" + env[i:j] + "}\n")
    oracle = [os.path.join(ROOT, "oracle", f) for f in sorted(os.listdir(os.path.join(ROOT, "oracle")))
              if f.endswith(".c")]
    exe = str(d / "headers_check")
    cmd = ["g++", "-O1", "-std=c++17", "-ffp-contract=off", "-fopenmp", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-I", str(d),
           os.path.join(ROOT, "tests", "host", "headers_check.cpp")] + oracle + ["-o", exe, "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


def test_device_headers_on_host(checker):
    r = subprocess.run([checker], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "0 failures" in r.stdout
