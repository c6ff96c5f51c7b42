"""The boundary from C: the header compiles as C99 on its own, and a plain C program built with
gcc against the header and librscape_hip.so scores an alignment through the C-ABI (GPU)."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from rscape_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
LIBDIR = os.path.join(ROOT, "r-scape_amd")


def test_header_compiles_as_c99(tmp_path):
    src = tmp_path / "only_header.c"
    src.write_text('#include "rscape_hip.h"\nint main(void) { rsg_hist h; (void)h; return RSG_OK; }\n')
    for std in ("c99", "c11"):
        subprocess.check_call(["gcc", "-std=" + std, "-Wall", "-Wextra", "-pedantic", "-Werror",
                               "-fsyntax-only", "-I", INC, str(src)])


def test_c_program_links_against_library(tmp_path):
    """gcc compiles and links the C caller (no GPU needed to link)."""
    exe = tmp_path / "abi_gtp"
    subprocess.check_call(["gcc", "-std=c99", "-Wall", "-Werror", "-I", INC,
                           os.path.join(ROOT, "tests", "c", "abi_gtp.c"), "-o", str(exe),
                           "-L", LIBDIR, "-lrscape_hip", "-Wl,-rpath," + LIBDIR])
    assert exe.exists()


@pytest.mark.gpu
def test_c_program_scores_gtp(tmp_path):
    """create -> rsg_cov_Calculate (GTp) -> rsg_corr_GetCOV from a C program, against the oracle."""
    exe = tmp_path / "abi_gtp"
    subprocess.check_call(["gcc", "-std=c99", "-O2", "-I", INC, os.path.join(ROOT, "tests", "c", "abi_gtp.c"),
                           "-o", str(exe), "-L", LIBDIR, "-lrscape_hip", "-Wl,-rpath," + LIBDIR])
    ax = synth.config(0, seed=11)["ax"]
    N, L = ax.shape
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        f.write(np.array([N, L], dtype=np.int64).tobytes())
        f.write(np.ascontiguousarray(ax, dtype=np.uint8).tobytes())
    print("C caller N=%d L=%d" % (N, L), flush=True)
    subprocess.check_call([str(exe), str(inp), str(out)], timeout=120)
    v = np.fromfile(out, dtype=np.float64)
    mn, mx, got = v[0], v[1], v[2:].reshape(L, L)
    pr = oracle.probs(ax)
    ref, _ = oracle.cov(pr, oracle.GT, oracle.CSELECT)
    ref, mm = oracle.corrected(ref, L, oracle.APC)
    ref = ref.reshape(L, L)
    iu = np.triu_indices(L, 1)
    floor = 1e-3 * (mm[1] - mm[0])
    assert np.max(np.abs(got[iu] - ref[iu]) / np.maximum(np.abs(ref[iu]), floor)) < 1e-10
    assert np.allclose([mn, mx], mm, rtol=1e-10)
    assert np.all(np.isneginf(np.diag(got)))
