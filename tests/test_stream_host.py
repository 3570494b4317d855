"""Host side of the streamed batch (wgpu-sigops_amd/csrc/stream_host.hpp: tile layout, per-chunk
bucketing and packing, unpacking), built for the CPU with g++ and run under AddressSanitizer and
UBSan against a stand-in for the kernel. The GPU tests cover the streamed batch end to end."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "wgpu-sigops_amd", "csrc")


def test_stream_pack_unpack(tmp_path):
    exe = str(tmp_path / "stream_host_test")
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-I" + CSRC, os.path.join(HERE, "cpp", "stream_host_test.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
