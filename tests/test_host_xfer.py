"""Host-side logic of the pinned-staging transfers (wgpu-sigops_amd/csrc/host_xfer.hpp), built for
the CPU with g++ and run under AddressSanitizer and UBSan: no GPU is involved. The GPU tests cover
the transfers themselves through every host entry point."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "wgpu-sigops_amd", "csrc")


def test_pool_and_packing(tmp_path):
    exe = str(tmp_path / "host_xfer_test")
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-D__HIP_PLATFORM_AMD__", "-I" + CSRC, "-I/opt/rocm/include",
                    os.path.join(HERE, "cpp", "host_xfer_test.cpp"), "-o", exe, "-pthread"], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
