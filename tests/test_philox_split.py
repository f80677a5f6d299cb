"""The per-packet Philox split of the batched walk (philox_packet / philox_packet_block in
pymonteopt_amd/csrc/philox.hpp) gives the plain Philox4x32-10 block bit for bit: the header
compiled for the host with g++ and run over 2e6 counters and keys, edge words included."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_packet_split_equals_philox(tmp_path):
    exe = tmp_path / "philox_split"
    subprocess.run(["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    "-I" + os.path.join(ROOT, "pymonteopt_amd", "csrc"),
                    os.path.join(ROOT, "tests", "c", "philox_split.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("0 of 2000000")
