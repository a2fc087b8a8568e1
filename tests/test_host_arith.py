"""CPU checks of the device arithmetic and the C ABI (no GPU needed).

The HIP kernels include via-rs_amd/csrc/modarith.h and ntt_lazy.h as
__host__ __device__ code; the same headers are compiled here with g++ and
checked against 128-bit integer arithmetic and the radix-2 NTT schedule.
"""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def build_and_run(src, tmp_path):
    exe = tmp_path / "check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-o", str(exe),
                    os.path.join(ROOT, "tests", "host", src)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("OK"), out.stdout
    return out.stdout


def test_modarith_against_u128(tmp_path):
    # exact fp64 modmul for every paper prime, the 256-bit gadget scale, the
    # u128 and int64 digit extractions and Garner reconstruction
    build_and_run("modarith_check.cpp", tmp_path)


def test_xcd_task_map(tmp_path):
    # workgroup -> (item, digit group, tile) map of the gadget MAC: a bijection
    # that keeps each client's tasks on the XCDs congruent to it
    build_and_run("xcd_map_check.cpp", tmp_path)


def test_lazy_ntt_passes(tmp_path):
    # radix-2^R passes over lazy residues: bit-identical canonical results
    # and the stated intermediate bounds, for every paper prime and schedule
    build_and_run("ntt_lazy_check.cpp", tmp_path)


def test_first_dim_limbs(tmp_path):
    # 7-bit limb split and the Horner fold of limb-product sums used by the
    # int8 matrix-core FirstDim: exact against 128-bit row sums
    build_and_run("fd_limbs_check.cpp", tmp_path)


def test_prg_blocks_match_shake256(tmp_path):
    # the device Keccak / Shake256Prg block against hashlib's SHAKE-256 of
    # seed || u64_le(k) (prg.rs:1-17), empty to 64-byte seeds, counters up
    # to 2^64 - 1
    import hashlib

    exe = tmp_path / "keccak"
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", str(exe), os.path.join(ROOT, "tests", "host", "keccak_check.cpp")],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60, check=True).stdout.splitlines()
    assert len(out) == 35
    for line in out:
        parts = line.split(" ")
        seed_hex, k, block = (parts[0], parts[1], parts[2]) if len(parts) == 3 else ("", parts[1], parts[2])
        want = hashlib.shake_256(bytes.fromhex(seed_hex) + int(k).to_bytes(8, "little")).digest(136).hex()
        assert block == want, (seed_hex, k)


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "via_gpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(via_[a-z_0-9]+)\s*\(", src, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib_path = os.path.join(ROOT, "via-rs_amd", "libvia_gpu.so")
    if not os.path.exists(lib_path):
        pytest.skip("libvia_gpu.so not built")
    lib = ctypes.CDLL(lib_path)
    names = declared_symbols()
    assert len(names) >= 20, names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_status_strings():
    lib_path = os.path.join(ROOT, "via-rs_amd", "libvia_gpu.so")
    if not os.path.exists(lib_path):
        pytest.skip("libvia_gpu.so not built")
    lib = ctypes.CDLL(lib_path)
    lib.via_status_str.restype = ctypes.c_char_p
    assert lib.via_status_str(2) == b"query length mismatch"
    assert lib.via_status_str(1) == b"dimension mismatch"
