"""CPU test: the oracle (oracle/*.c) and the host build of the device headers (tests/devcheck) built
with AddressSanitizer and UBSan (every report fatal), then run through their own tests --
tests/test_oracle.py, tests/test_oracle_pins.py and tests/test_devcheck.py -- in a child process
that preloads the ASan runtime. No GPU is involved."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
SAN = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=all"]


def _runtime():
    r = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True)
    path = r.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


@pytest.mark.skipif(_runtime() is None, reason="gcc's AddressSanitizer runtime not found")
def test_oracle_and_devcheck_under_asan_ubsan(tmp_path):
    oracle = os.path.join(REPO, "oracle")
    osan = str(tmp_path / "liboracle_san.so")
    dsan = str(tmp_path / "libdevcheck_san.so")
    srcs = [os.path.join(oracle, f) for f in ("bn.c", "sha2.c", "secp.c", "ed25519.c", "gen.c")]
    subprocess.run(["gcc", "-std=gnu11", "-fPIC", "-shared", "-pthread"] + SAN + srcs + ["-o", osan, "-lm"],
                   check=True)
    subprocess.run(["g++", "-std=c++17", "-fPIC", "-shared", "-Wno-unknown-pragmas", "-DSIGOPS_HOST_CHECK=1"] + SAN +
                   ["-I" + os.path.join(REPO, "wgpu-sigops_amd", "csrc"), os.path.join(HERE, "devcheck", "devcheck.cpp"),
                    "-o", dsan], check=True)
    env = dict(os.environ, LD_PRELOAD=_runtime(), ASAN_OPTIONS="detect_leaks=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", SIGOPS_ORACLE_SO=osan, SIGOPS_DEVCHECK_SO=dsan)
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-s", "-p", "no:cacheprovider", "-m", "not gpu",
                        os.path.join(HERE, "test_oracle.py"), os.path.join(HERE, "test_oracle_pins.py"),
                        os.path.join(HERE, "test_devcheck.py")],
                       env=env, capture_output=True, text=True, timeout=900, cwd=REPO)
    tail = (r.stdout + r.stderr)[-4000:]
    assert r.returncode == 0, tail
    assert "passed" in r.stdout and "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, tail
