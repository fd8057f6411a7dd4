"""The checkpoint image layout (lh_checkpoint_header in include/lenia_hip.h) as the
Python reader sees it, pinned against the C compiler's own layout of the header, and
the reader on an image assembled by hand.  No GPU."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from lenia_amd import _lib as L
from lenia_amd.engine import read_checkpoint

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROBE = r"""
#include <stdio.h>
#include <stddef.h>
#include "lenia_hip.h"
int main(void) {
    printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(lh_checkpoint_header),
           offsetof(lh_checkpoint_header, version), offsetof(lh_checkpoint_header, conv),
           offsetof(lh_checkpoint_header, shape), offsetof(lh_checkpoint_header, steps),
           offsetof(lh_checkpoint_header, payload_bytes), sizeof(lh_params), sizeof(lh_kernel),
           (size_t)LH_CHECKPOINT_VERSION);
    return 0;
}
"""


def test_layout_matches_the_c_header(tmp_path):
    src = tmp_path / "probe.c"
    src.write_text(PROBE)
    exe = tmp_path / "probe"
    try:
        subprocess.run(["gcc", "-I", os.path.join(REPO, "include"), str(src), "-o", str(exe)], check=True,
                       capture_output=True)
    except (OSError, subprocess.CalledProcessError) as e:
        pytest.skip(f"no C compiler: {e}")
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    H = L.CheckpointHeader
    want = [C.sizeof(H), H.version.offset, H.conv.offset, H.shape.offset, H.steps.offset, H.payload_bytes.offset,
            C.sizeof(L.Params), C.sizeof(L.Kernel), L.CHECKPOINT_VERSION]
    assert got == want
    assert got[0] == 64


def _image(W=4, H=6, D=5, Cn=2, K=3, steps=17, conv=0, seed=0):
    rng = np.random.default_rng(seed)
    N = W * H * D
    p = L.Params()
    p.dt, p.sigma, p.quirks = 0.2, 0.65, 0x155
    ks = (L.Kernel * K)()
    for i in range(K):
        ks[i].c0, ks[i].c1, ks[i].radius, ks[i].m = i % Cn, (i + 1) % Cn, 3 + i, 0.1 * i
    fields = {"A": rng.random(Cn * N, dtype=np.float32), "GA": rng.random(Cn * N, dtype=np.float32),
              "S": rng.random(N, dtype=np.float32), "V": rng.standard_normal(3 * Cn * N).astype(np.float32)}
    payload = b"".join(fields[k].astype("<f4").tobytes() for k in ("A", "GA", "S", "V"))
    h = L.CheckpointHeader()
    h.magic, h.version, h.header_bytes = b"LHCK", 1, C.sizeof(L.CheckpointHeader)
    h.params_bytes, h.kernel_bytes, h.conv = C.sizeof(L.Params), C.sizeof(L.Kernel), conv
    h.shape = L.Shape(W, H, D, Cn, K)
    h.steps, h.payload_bytes = steps, len(payload)
    return bytes(h) + bytes(p) + bytes(ks) + payload, fields, p, ks


def test_reader_recovers_every_part():
    img, fields, p, ks = _image()
    ck = read_checkpoint(img)
    assert ck["steps"] == 17 and ck["conv"] == 0 and ck["version"] == 1
    s = ck["shape"]
    assert (s.W, s.H, s.D, s.C, s.K) == (4, 6, 5, 2, 3)
    assert bytes(ck["params"]) == bytes(p) and bytes(ck["kernels"]) == bytes(ks)
    for k, v in fields.items():
        assert np.array_equal(ck["fields"][k].view(np.uint32), v.view(np.uint32)), k


@pytest.mark.parametrize("damage", ["magic", "version", "truncated", "extra", "payload"])
def test_reader_rejects_damaged_images(damage):
    img, *_ = _image()
    b = bytearray(img)
    if damage == "magic":
        b[0:4] = b"LHCX"
    elif damage == "version":
        b[4] = 2
    elif damage == "truncated":
        b = b[:-4]
    elif damage == "extra":
        b += b"\0\0\0\0"
    elif damage == "payload":
        b[56] ^= 4  # payload_bytes
    with pytest.raises(ValueError):
        read_checkpoint(bytes(b))
