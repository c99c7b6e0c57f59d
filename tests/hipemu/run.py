"""Bounds check of the HIP library on a machine without a GPU.

Builds mdspherocyl_amd/csrc with the CPU emulation of the HIP runtime in
this folder under AddressSanitizer (ROCm's host clang), then runs the given pytest arguments
(normally the GPU parity tests) against that library instead of the gfx950
build.  Every kernel runs block by block on host threads, so out-of-bounds
device accesses are reported by ASan with the kernel's source line.

    python tests/hipemu/run.py tests/test_gpu_parity.py -m gpu -x -v

The script re-executes itself under the ASan runtime (LD_PRELOAD) before
anything else is imported; it never touches a GPU.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "_build", "libmsc_emu.so")
# the host clang of ROCm: the kernels' _Float16 vectors need it
CLANG = "/opt/rocm/lib/llvm/bin/clang++"


def build():
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    src = os.path.join(ROOT, "mdspherocyl_amd", "csrc")
    cmd = [CLANG, "-std=c++20", "-O1", "-g", "-fsanitize=address", "-shared-libasan", "-fno-omit-frame-pointer",
           "-fPIC", "-shared", "-I", HERE, "-x", "c++", os.path.join(src, "msc_kernels.hip"),
           os.path.join(src, "msc_api.hip"), os.path.join(src, "msc_conf.cpp"), os.path.join(src, "msc_place.cpp"),
           "-o", OUT, "-lpthread"]
    subprocess.run(cmd, check=True)


def main():
    if os.environ.get("MSC_EMU_CHILD") != "1":
        build()
        asan = subprocess.run([CLANG, "-print-file-name=libclang_rt.asan-x86_64.so"], capture_output=True,
                              text=True, check=True).stdout.strip()
        env = dict(os.environ, MSC_EMU_CHILD="1", LD_PRELOAD=asan,
                   ASAN_OPTIONS="detect_leaks=0:abort_on_error=1")
        sys.exit(subprocess.run([sys.executable, os.path.abspath(__file__)] + sys.argv[1:], env=env).returncode)
    sys.path.insert(0, ROOT)
    import mdspherocyl_amd._lib as lib
    lib.LIBPATH = OUT
    import pytest
    sys.exit(pytest.main(sys.argv[1:]))


if __name__ == "__main__":
    main()
