"""A plain-C program calls the C ABI in the order QEX's shim would
(tests/c/abi_sequence.c): compiled here with gcc, run on the GPU."""
import os
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _build(tmp):
    exe = os.path.join(tmp, "abi_sequence")
    libdir = os.path.join(ROOT, "qex_amd")
    subprocess.run(["gcc", "-O2", "-std=gnu11", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "c", "abi_sequence.c"), "-o", exe,
                    "-L" + libdir, "-lqexhip", "-Wl,-rpath," + libdir, "-lm"], check=True)
    return exe


def test_c_program_compiles_against_header():
    with tempfile.TemporaryDirectory() as tmp:
        assert os.path.exists(_build(tmp))


@pytest.mark.gpu
def test_c_abi_sequence_runs():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    with tempfile.TemporaryDirectory() as tmp:
        exe = _build(tmp)
        p = subprocess.run([exe], capture_output=True, text=True, timeout=120)
        print(p.stdout, p.stderr)
        assert p.returncode == 0 and p.stdout.strip().endswith("OK")
