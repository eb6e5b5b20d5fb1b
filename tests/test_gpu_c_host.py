"""GPU: a plain C program (tests/c_host/dop_host.c) drives the library through the C ABI
only -- keygen, key upload, a DOp stream with SYNC and ack, and a radix ADD through the
serving engine -- and checks the decrypted results."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def test_c_host_program():
    exe = os.path.join(HERE, "c_host", "dop_host")
    if not os.path.exists(exe):
        pytest.fail("tests/c_host/dop_host is not built (run __graft_entry__.build())")
    print("case: C host program (DOp stream + server ADD on 8 blocks)", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "DOp stream: PBS(2 + 3) = 10 (expected 10)" in r.stdout
    assert "server: ADD 40000 + 23456 = 63456" in r.stdout
